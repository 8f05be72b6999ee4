/* CPU oracle for the LBVH build / refit / traversal / thin-lens render path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a double-precision restatement of the
 * reference algorithm (mbartling/cuda-rt) used as the checker for the HIP
 * library and as the CPU baseline in bench.py. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product library never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates. Index
 * conventions follow inc/bvh.cuh:25-28: internal nodes are [0, n-2],
 * leaves are [n-1, 2n-2], a leaf has left == right == triangle id.
 */
#ifndef RT_ORACLE_H
#define RT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t n;                    /* triangles */
  int32_t nv;                   /* vertices */
  const float *vx, *vy, *vz;    /* vertex positions, nv each */
  const float *nx, *ny, *nz;    /* per-vertex shading normals, nv each */
  const int32_t *idx;           /* 3n vertex indices */
  const int32_t *mat_id;        /* n material ids, or NULL for material 0 */
  const double *pos64;          /* optional: nv positions as doubles (xyz), used instead of vx/vy/vz */
} orc_mesh;

typedef struct {
  int32_t n;                    /* triangles (leaves) */
  const int32_t *parent;        /* 2n-1; root's parent is -1 */
  const int32_t *left;          /* 2n-1 */
  const int32_t *right;         /* 2n-1 */
  const float *lo;              /* (2n-1)*3 node bounds, AoS xyz */
  const float *hi;              /* (2n-1)*3 */
} orc_bvh;

typedef struct {
  double ka[3], kd[3], ks[3], kt[3], ke[3];
  double shininess;
} orc_material;

typedef struct {
  double position[3];
  double color[3];
  double radius;
  double c0, c1, c2;            /* falloff 1/(c0 + c1 d + c2 d^2) */
} orc_light;

typedef struct {
  int32_t width, height;
  double fov_deg;
  double fstop;                 /* aperture = 1/fstop */
  double focal_point;
  double eye[3];
} orc_camera;

enum {
  ORC_RANDVEC_RTL = 1,          /* evaluate randVec3d's three draws z,y,x */
  ORC_CLAMP = 2,                /* clamp the averaged colour to [0,1] */
  /* Planted mistakes, for tests that show a statistic can see them; never
   * set outside tests. */
  ORC_PLANT_LIGHT_OFFSET = 1 << 24, /* light jitter (rand - 0.5) * radius instead of rand * radius - 0.5 */
  ORC_PLANT_APERTURE = 1 << 25,     /* lens aperture 1.25 / fstop */
  ORC_PLANT_LENS_RADIUS = 1 << 26   /* lens radius aperture * u / 2 (no square root) */
};

typedef struct {
  int32_t spp;
  int32_t depth;                /* 0: reference shading only; >0 diffuse bounces */
  uint32_t seed;                /* 0 = reference seeding */
  int32_t flags;
} orc_render_opts;

/* A1: per-triangle AABB (kernels/scene.cu:88-120). lo/hi are n*3. */
void orc_tri_aabb(const orc_mesh *m, float *lo, float *hi);
/* A2: scene bounds (kernels/scene.cu:72-85), init +-1e9, fminf/fmaxf. */
void orc_scene_bounds(int32_t n, const float *lo, const float *hi, double *mn, double *mx);
/* A3: Morton codes of AABB centroids (kernels/bvh.cu:13-61). */
void orc_morton30(int32_t n, const float *lo, const float *hi, const double *mn, const double *mx, uint32_t *codes);
void orc_morton63(int32_t n, const float *lo, const float *hi, const double *mn, const double *mx, uint64_t *codes);
uint32_t orc_expand_bits10(uint32_t v);
/* A4: stable sort by key (kernels/bvh.cu:267-273); ids start as identity. */
void orc_sort30(int32_t n, uint32_t *codes, uint32_t *ids);
void orc_sort63(int32_t n, uint64_t *codes, uint32_t *ids);
/* A5/A6: Karras hierarchy (kernels/bvh.cu:71-207). Arrays are 2n-1. */
void orc_hierarchy30(int32_t n, const uint32_t *codes, const uint32_t *ids, int32_t *parent, int32_t *left, int32_t *right);
void orc_hierarchy63(int32_t n, const uint64_t *codes, const uint32_t *ids, int32_t *parent, int32_t *left, int32_t *right);
/* Brute-force top-down radix tree (Karras 2012 definition) for pinning A5/A6. */
void orc_topdown30(int32_t n, const uint32_t *codes, const uint32_t *ids, int32_t *parent, int32_t *left, int32_t *right);
void orc_topdown63(int32_t n, const uint64_t *codes, const uint32_t *ids, int32_t *parent, int32_t *left, int32_t *right);
/* A7: bottom-up refit with flags (kernels/bvh.cu:209-258), leaves visited
 * in the order given by perm (NULL = ascending); node lo/hi are (2n-1)*3 and
 * tri_lo/tri_hi are the per-triangle boxes. */
void orc_refit(int32_t n, const int32_t *parent, const int32_t *left, const int32_t *right,
               const float *tri_lo, const float *tri_hi, const int32_t *perm, float *lo, float *hi);
/* Brute-force subtree bounds: min/max over every triangle below each node. */
void orc_refit_brute(int32_t n, const int32_t *left, const int32_t *right,
                     const float *tri_lo, const float *tri_hi, float *lo, float *hi);
/* Full build = A1..A7. codes is n*8 bytes (uint64), written sorted. */
void orc_build(const orc_mesh *m, int32_t bits, uint64_t *codes, uint32_t *ids, int32_t *parent,
               int32_t *left, int32_t *right, float *lo, float *hi, double *mn, double *mx, int32_t nthreads);

/* A9-A12: traversal (kernels/bvh.cu:297-449, inc/bbox.h:82-102).
 * Rays are AoS xyz doubles. Outputs: hit triangle (-1 = miss), t, alpha,
 * beta, shading normal (3 per ray), node-visit and triangle-test counts. */
void orc_trace(const orc_mesh *m, const orc_bvh *b, int64_t nrays, const double *orig, const double *dir,
               int32_t any_hit, int32_t *tri, double *t, double *alpha, double *beta, double *normal,
               int64_t *visits, int32_t nthreads);
/* Brute force closest hit: min t over all triangles passing the same
 * triangle test, ties broken by smallest sorted leaf position. leafpos[tri]
 * gives the sorted position of each triangle. */
void orc_trace_brute(const orc_mesh *m, const int32_t *leafpos, int64_t nrays, const double *orig,
                     const double *dir, int32_t *tri, double *t);
/* A10 alone: one ray, one triangle. Returns 1 on hit. */
int32_t orc_intersect_triangle(const orc_mesh *m, int32_t tri, const double *orig, const double *dir,
                               double *t, double *alpha, double *beta, double *normal);
/* A9 alone. */
int32_t orc_box_hit(const float *lo, const float *hi, const double *orig, const double *dir);

/* A14: mrand (kernels/mrand.cu:3-41). */
uint32_t orc_wang_hash(uint32_t s);
uint32_t orc_xorshift(uint32_t *state);
double orc_rand_double(uint32_t *state);
/* Seed index of pixel (px,py): the 16x16-block linear id (inc/indexing.h:4-9). */
uint32_t orc_seed_index(int32_t px, int32_t py, int32_t width);
void orc_sincos(double x, double *s, double *c);
/* Cosine-weighted direction about the unit normal n from two uniforms, in the
 * frame of src/ray.cpp:5-32 (the diffuse-bounce extension). */
void orc_cosine_dir(const double *n, double u1, double u2, double *dir);

/* A15: primary ray of sample `sample` of launch pixel (px,py) with the
 * reference's draw order (kernels/raytracer.cu:126-154). */
void orc_primary_ray(const orc_camera *cam, const orc_render_opts *opts, int32_t px, int32_t py, int32_t sample,
                     double *orig, double *dir);

/* A15-A18: render (kernels/raytracer.cu:94-249, kernels/material.cu:6-40,
 * kernels/Light.cu:5-86). rgb is width*height*3 floats, row-flipped like
 * the reference image. If pixels != NULL only those launch pixels
 * (py*width+px) are rendered. counts[0..2] += primary, shadow, bounce rays. */
void orc_render(const orc_mesh *m, const orc_bvh *b, const orc_material *mats, int32_t nmat,
                const orc_light *light, const orc_camera *cam, const orc_render_opts *opts,
                float *rgb, const int64_t *pixels, int64_t npix, int64_t *counts, int32_t nthreads);
/* orc_render with optional outputs (each may be NULL): rgb64 the same
 * image in doubles, records per pixel (py*width+px), sample and bounce
 * (depth+1 slots) a pair (hit triangle or -1, occluder: -1 none, -2 opaque,
 * else the transmitting occluder's triangle), branches per slot the
 * shading's branch bits (bit 0: N.L > 0, bit 1: falloff clamped, bit 2:
 * normal flipped for the bounce, bits 3-4: bounce frame axis); unused
 * slots hold -1. */
void orc_render64(const orc_mesh *m, const orc_bvh *b, const orc_material *mats, int32_t nmat,
                  const orc_light *light, const orc_camera *cam, const orc_render_opts *opts, float *rgb,
                  double *rgb64, int32_t *records, int32_t *branches, const int64_t *pixels, int64_t npix,
                  int64_t *counts, int32_t nthreads);

#ifdef __cplusplus
}
#endif
#endif
