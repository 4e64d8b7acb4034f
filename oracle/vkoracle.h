/*
 * vkoracle.h - CPU oracle for the ray-intersection path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference's algorithms, used
 * as the comparison reference by tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py.
 * Nothing in the product (libvkrt.so, the C++ host layer) links, loads or calls it.
 *
 * Built with -O3 -ffp-contract=off on x86-64 (SSE2 scalar float, no FMA, no FTZ/DAZ), which is
 * the arithmetic of the reference's Release build of src/intersection.cpp (CMakeLists.txt:51-68).
 *
 * Ray layout: 8 floats {ox, oy, oz, tmin, dx, dy, dz, tmax}. Hit layout: {t, u, v, prim(u32)}.
 */
#ifndef VKORACLE_H
#define VKORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* intersect(const Triangle&, const Ray&), src/intersection.cpp:114-129. Returns 0 (REJECT) or 1
 * (INTERSECT); t/u/v are written only on INTERSECT. tri = 9 floats, o/d = 3 floats. */
int orc_intersect_tri(const float *tri, const float *o, const float *d, float *t, float *u, float *v);
/* the same algorithm evaluated in double precision from the same float inputs (FP64 twin) */
int orc_intersect_tri_f64(const float *tri, const float *o, const float *d, double *t, double *u, double *v);

/* intersect(const OBB&, const Ray&), src/intersection.cpp:152-184; obb = 15 floats
 * {center, axis0, axis1, axis2, half_lengths}. Returns type, writes distance. */
int orc_intersect_obb(const float *obb, const float *o, const float *d, float *dist);
/* AABB::intersect(const Ray&), src/intersection.cpp:308-312 (via OBB(aabb, mat4(1)), :219-227) */
int orc_intersect_aabb(const float *box6, const float *o, const float *d, float *dist);
/* intersect(const Sphere&, const Ray&), src/intersection.cpp:133-148 */
int orc_intersect_sphere(const float *sphere4, const float *o, const float *d, float *dist);
/* intersect(const Plane&, const Ray&), src/intersection.cpp:10-21 */
int orc_intersect_plane(const float *plane4, const float *o, const float *d, float *dist);
/* compute_aabb, src/intersection.cpp:96-106 */
void orc_compute_aabb(const float *points, uint32_t n, float *out6);
void orc_bounding_sphere(const float *points, uint32_t n, float *out4);
void orc_centroid(const float *points, uint32_t n, float *out3);
void orc_transmittance(const float *tris, const float *trans, const float *color, const uint8_t *two_sided,
                       uint32_t n_tri, const float *rays8, uint32_t n_ray, int cull, float *out4);
void orc_scene_closest(const float *tris, const uint32_t *tri_offset, const float *w2o, const uint32_t *inst_mesh,
                       const uint32_t *inst_custom, uint32_t n_inst, const float *rays8, uint32_t n_ray,
                       uint32_t *out8);

/* brute-force closest / any hit over flat triangles (9 floats each), split over nthreads */
void orc_brute_closest(const float *tris, uint32_t n_tri, const float *rays8, uint32_t n_ray, void *hits16,
                       int nthreads);
void orc_brute_any(const float *tris, uint32_t n_tri, const float *rays8, uint32_t n_ray, uint8_t *occluded,
                   int nthreads);
/* brute-force closest hit plus, per ray, the library traversal's culling criterion for the winner
 * (orc_cull_safe; 1 for a miss) */
void orc_brute_closest_cull_safe(const float *tris, uint32_t n_tri, const float *rays8, uint32_t n_ray,
                                 void *hits16, uint8_t *safe, int nthreads);
/* 1 when the library's padded box of triangle `tri` is entered by the ray with best t = t (see vkoracle.c) */
int orc_cull_safe(const float *tri, const float *ray8, float t);
/* closest hit of one ray, evaluated in FP64 (same acceptance rule); returns prim or 0xFFFFFFFF */
uint32_t orc_closest_f64(const float *tris, uint32_t n_tri, const float *ray8, double *t, double *u, double *v);

/* Geometry::IcoSphere, src/Geometry.cpp:566-613 with tessellate :141-192 */
void orc_icosphere_size(uint32_t subdiv, uint32_t *n_vert, uint32_t *n_idx);
void orc_icosphere(float radius, uint32_t subdiv, float *pos, uint32_t *idx);
/* Geometry::Plane(w, h, sw, sh, xy_plane = false), src/Geometry.cpp:313-371 */
void orc_plane(float width, float height, uint32_t sw, uint32_t sh, float *pos, uint32_t *idx);

/* RNG utilities of the path tracer */
/* xxhash32(uint2) shaders/slang/utils/hash.slang:9-18 */
uint32_t orc_xxhash32(uint32_t a, uint32_t b);
/* lcg / rnd, shaders/slang/utils/random.slang:27-36 */
uint32_t orc_lcg(uint32_t *state);
float orc_rnd(uint32_t *state);
/* Hammersley point (i, n), shaders/slang/utils/sampling.slang:9-13 */
void orc_hammersley(uint32_t i, uint32_t n, float *out2);

#ifdef __cplusplus
}
#endif

#endif
