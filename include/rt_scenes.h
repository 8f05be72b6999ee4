/* Deterministic synthetic scene generators for the benchmark and test
 * workloads (host code, no GPU needed). The reference loads OBJ files
 * (src/scene.cpp:6-58); these generators have no counterpart there and
 * produce the procedural scenes the workloads are defined on. All random
 * draws come from splitmix64 keyed by (seed, element, draw), so output is
 * identical for any thread count. Arrays are caller-allocated host memory.
 */
#ifndef RT_SCENES_H
#define RT_SCENES_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Random triangle soup: centroids uniform in [0,1)^3 * scale + offset,
 * equilateral triangles with edge lognormal(ln(edge_median), edge_sigma),
 * uniformly random orientation (Shoemake quaternion). nv = 3n; each
 * triangle owns its 3 vertices; normals are the face normal. */
void rt_gen_soup(int32_t n, uint64_t seed, double edge_median, double edge_sigma, const double offset[3],
                 double scale, float *vx, float *vy, float *vz, float *nx, float *ny, float *nz, int32_t *idx);

/* Geodesic icosphere of the given frequency: 20 f^2 triangles and
 * 10 f^2 + 2 shared vertices. If noise_amp != 0 the radius is multiplied
 * by 1 + noise_amp * n(p), n = 3-octave value noise on a lattice seeded by
 * noise_seed (frequencies 4, 8, 16; amplitudes 1, 1/2, 1/4, normalised to
 * [-1, 1]). Normals are area-weighted vertex normals. */
void rt_gen_icosphere_counts(int32_t freq, int32_t *nv, int32_t *n);
void rt_gen_icosphere(int32_t freq, double radius, const double center[3], double noise_amp, uint64_t noise_seed,
                      float *vx, float *vy, float *vz, float *nx, float *ny, float *nz, int32_t *idx);

/* Gaussian vertex jitter (sigma) for refit frames; out may alias in. */
void rt_gen_jitter(int32_t nv, const float *vx, const float *vy, const float *vz, double sigma, uint64_t seed,
                   float *ox, float *oy, float *oz);

#ifdef __cplusplus
}
#endif
#endif
