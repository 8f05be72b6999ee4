/* Scene input and image output for the ray-tracer core (host side).
 *
 * rt_obj_*: the OBJ/MTL path of Scene_h::LoadObj (src/scene.cpp:6-58) over
 * the vendored tinyobj loader (inc/tiny_obj_loader.h, LoadObj with
 * triangulation on). The result is the mesh layout of rt_mesh_view: each
 * distinct (position index, normal index) pair of the file becomes one
 * vertex, so positions and normals share the triangle indices while every
 * triangle keeps the reference's position and normal. Positions and normals
 * are parsed to double as the reference does and stored as fp32 (the
 * reference's AABBs and Morton codes round them to float as well,
 * inc/vec.h:158-165); the parsed doubles stay available.
 *
 * rt_write_bmp: RayTracer::pullRaytracedImage + writeBMP
 * (inc/raytracer.h:50-62, src/bitmap.cpp:100-155).
 */
#ifndef RT_IO_H
#define RT_IO_H
#include <stddef.h>
#include <stdint.h>

#include "rt_api.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rt_obj rt_obj;

/* Parse `path` and the MTL files it names (looked up under mtl_basepath,
 * as tinyobj's MaterialFileReader does, inc/tiny_obj_loader.h:1002-1025).
 * Faces are fan-triangulated in file order; the face grouping of tinyobj
 * (shapes flushed at g / o / usemtl) is followed, including its rule that
 * a group flushed with no pending faces is dropped. Material ids are the
 * MTL order (-1: no usemtl or an unknown name). On failure *out is NULL and
 * err (if errlen > 0) holds a message. */
rt_status rt_obj_load(const char *path, const char *mtl_basepath, rt_obj **out, char *err, size_t errlen);
/* Counts: unique vertices, triangles, materials, shapes, and the file's raw
 * position and normal counts (the reference's numVertices / numNormals). */
void rt_obj_counts(const rt_obj *obj, int32_t *nv, int32_t *n, int32_t *nmat, int32_t *nshapes, int32_t *raw_v,
                   int32_t *raw_vn);
/* Host arrays: nv positions and normals (fp32; a missing normal is 0),
 * 3n vertex indices, n material ids. Any pointer may be NULL. */
void rt_obj_mesh(const rt_obj *obj, float *vx, float *vy, float *vz, float *nx, float *ny, float *nz, int32_t *idx,
                 int32_t *mat_id);
/* The nv positions as the parsed doubles (xyz interleaved). */
void rt_obj_positions64(const rt_obj *obj, double *xyz);
/* nmat materials in the reference's Material fields (src/scene.cpp:22-35):
 * ka = ambient, kd = diffuse, ks = specular, kt = transmittance,
 * ke = emission, shininess = Ns. */
void rt_obj_materials(const rt_obj *obj, rt_material *mats);
void rt_obj_free(rt_obj *obj);

/* 24-bit BMP of an image laid out as rt_render writes it (width*height*3
 * floats, row-flipped). Each channel is clamped to [0, 1] and quantised as
 * (unsigned char)(255.0 * v) (kernels/scene.cu:137, inc/raytracer.h:58-60),
 * written bottom-up in BGR order with rows padded to 4 bytes (the pad
 * bytes are zero). */
rt_status rt_write_bmp(const char *path, int32_t width, int32_t height, const float *rgb);

#ifdef __cplusplus
}
#endif
#endif
