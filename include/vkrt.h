/*
 * vkrt.h - C ABI of the MI355X (gfx950) ray-intersection and path-tracing core.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns an int status
 * (VKRT_OK == 0, negative on error) and never lets a C++ exception cross the boundary.
 * The message of the last failing call on a context is returned by vkrt_last_error().
 *
 * Pointer residency: array arguments are host pointers unless VKRT_DEVICE_PTRS is set in
 * `flags`, in which case every array argument of that call is a device (HBM) pointer.
 * Calls are synchronous on the context's stream unless VKRT_ASYNC is set (then call
 * vkrt_sync before reading results). A context is single-threaded; separate contexts may
 * be used concurrently from separate host threads.
 *
 * Interface correspondence with crocdialer/vierkant (reference file:line):
 *   vkrt_ray / vkrt_hit            <- struct Ray, ray_triangle_intersection
 *                                     (include/vierkant/intersection.hpp:145-170, :18-51)
 *   vkrt_intersect_triangles       <- intersect(const Triangle&, const Ray&)      (src/intersection.cpp:114-129)
 *   vkrt_intersect_aabbs           <- AABB::intersect(const Ray&)                 (src/intersection.cpp:308-312)
 *   vkrt_intersect_obbs            <- intersect(const OBB&, const Ray&)           (src/intersection.cpp:152-184)
 *   vkrt_intersect_spheres         <- intersect(const Sphere&, const Ray&)        (src/intersection.cpp:133-148)
 *   vkrt_intersect_planes          <- intersect(const Plane&, const Ray&)         (src/intersection.cpp:10-21)
 *   vkrt_compute_aabb              <- compute_aabb(const std::vector<vec3>&)      (src/intersection.cpp:96-106)
 *   vkrt_intersect_brute_*         <- a closest/any-hit loop over intersect(Triangle, Ray); the reference
 *                                     has no such loop, its callers test one triangle at a time
 *   vkrt_bvh_build / _refit        <- RayBuilder::create_mesh_structures / build_scene_acceleration
 *                                     (src/RayBuilder.cpp:98-303, :783-1035)
 *   vkrt_trace_closest / _any      <- TraceRay closest-hit / ACCEPT_FIRST_HIT shadow queries
 *                                     (shaders/slang/raypipeline.slang:182, ray/visibility_test.slang:45-87)
 *   vkrt_icosphere / vkrt_plane    <- Geometry::IcoSphere / Geometry::Plane       (src/Geometry.cpp:566-613, :313-371)
 */
#ifndef VKRT_H
#define VKRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VKRT_VERSION_MAJOR 0
#define VKRT_VERSION_MINOR 1

/* status codes */
enum
{
    VKRT_OK = 0,
    VKRT_E_INVALID_ARG = -1,
    VKRT_E_OOM = -2,
    VKRT_E_HIP = -3,
    VKRT_E_UNSUPPORTED = -4,
    VKRT_E_NO_DEVICE = -5,
    VKRT_E_INTERNAL = -6
};

/* call flags */
enum
{
    VKRT_HOST_PTRS = 0,
    VKRT_DEVICE_PTRS = 1,
    VKRT_ASYNC = 2
};

/* intersection_type (include/vierkant/intersection.hpp:18-23) */
enum
{
    VKRT_REJECT = 0,
    VKRT_INTERSECT = 1,
    VKRT_INSIDE = 2
};

#define VKRT_INVALID_PRIM 0xFFFFFFFFu

typedef struct vkrt_ctx vkrt_ctx;
typedef struct vkrt_bvh vkrt_bvh;

/* A ray: origin, valid-interval start, direction, valid-interval end (32 B, 16-B aligned).
 * A candidate hit at distance t counts iff tmin < t < tmax (both strict). The direction is
 * used as given (the reference normalizes in Ray's constructor on the host). */
typedef struct vkrt_ray
{
    float ox, oy, oz, tmin;
    float dx, dy, dz, tmax;
} vkrt_ray;

/* Closest-hit record (16 B). prim == VKRT_INVALID_PRIM on a miss, then t = +inf, u = v = 0. */
typedef struct vkrt_hit
{
    float t, u, v;
    uint32_t prim;
} vkrt_hit;

/* Result of one ray-triangle test, ray_triangle_intersection (intersection.hpp:44-51). */
typedef struct vkrt_tri_result
{
    uint32_t type;
    float distance, u, v;
} vkrt_tri_result;

/* Result of one ray-volume test, ray_intersection (intersection.hpp:28-38). */
typedef struct vkrt_result
{
    uint32_t type;
    float distance;
} vkrt_result;

/* ---------------------------------------------------------------------------------------------
 * context
 * ------------------------------------------------------------------------------------------- */

/* Create a context on HIP device `device`. `hip_stream` is an existing hipStream_t to run on,
 * or NULL to let the context create its own. */
int vkrt_create(int device, void *hip_stream, vkrt_ctx **out);
void vkrt_destroy(vkrt_ctx *ctx);
const char *vkrt_last_error(const vkrt_ctx *ctx);
int vkrt_sync(vkrt_ctx *ctx);
/* the hipStream_t the context launches on */
void *vkrt_stream(const vkrt_ctx *ctx);
int vkrt_version(void);
/* number of HIP devices visible (0 when none); never initialises a context */
int vkrt_device_count(void);

/* Device memory helpers for callers without their own HIP allocator (no counterpart in the reference,
 * whose buffers are Vulkan/VMA allocations). vkrt_copy kinds: 0 host->device, 1 device->host,
 * 2 device->device; synchronous on the context's stream unless VKRT_ASYNC is in flags. */
int vkrt_device_alloc(vkrt_ctx *ctx, size_t bytes, void **out);
int vkrt_device_free(vkrt_ctx *ctx, void *ptr);
int vkrt_copy(vkrt_ctx *ctx, void *dst, const void *src, size_t bytes, int kind, int flags);

/* ---------------------------------------------------------------------------------------------
 * pairwise intersection tests: element i tests primitive i against ray i (n pairs)
 * ------------------------------------------------------------------------------------------- */

/* tris: 9 floats per triangle (v0, v1, v2) */
int vkrt_intersect_triangles(vkrt_ctx *ctx, const float *tris, const vkrt_ray *rays, uint32_t n,
                             vkrt_tri_result *out, int flags);
/* boxes: 6 floats per AABB (min.xyz, max.xyz) */
int vkrt_intersect_aabbs(vkrt_ctx *ctx, const float *boxes, const vkrt_ray *rays, uint32_t n, vkrt_result *out,
                         int flags);
/* obbs: 15 floats per OBB (center.xyz, axis[0].xyz, axis[1].xyz, axis[2].xyz, half_lengths.xyz) */
int vkrt_intersect_obbs(vkrt_ctx *ctx, const float *obbs, const vkrt_ray *rays, uint32_t n, vkrt_result *out,
                        int flags);
/* spheres: 4 floats per sphere (center.xyz, radius) */
int vkrt_intersect_spheres(vkrt_ctx *ctx, const float *spheres, const vkrt_ray *rays, uint32_t n,
                           vkrt_result *out, int flags);
/* planes: 4 floats per plane (coefficients a, b, c, d with a unit normal) */
int vkrt_intersect_planes(vkrt_ctx *ctx, const float *planes, const vkrt_ray *rays, uint32_t n, vkrt_result *out,
                          int flags);
/* AABB of n points (3 floats each); out: 6 floats (min.xyz, max.xyz). n == 0 gives the
 * reference's empty box (min = FLT_MAX, max = -FLT_MAX). */
int vkrt_compute_aabb(vkrt_ctx *ctx, const float *points, uint32_t n, float *out6, int flags);

/* ---------------------------------------------------------------------------------------------
 * brute force: every ray against every triangle
 * ------------------------------------------------------------------------------------------- */

/* Closest hit: smallest t with tmin < t < tmax among INTERSECT results; equal t resolves to the
 * lowest triangle index. */
int vkrt_intersect_brute_closest(vkrt_ctx *ctx, const float *tris, uint32_t n_tri, const vkrt_ray *rays,
                                 uint32_t n_ray, vkrt_hit *out, int flags);
/* Any hit: occluded[i] = 1 iff some triangle gives INTERSECT with tmin < t < tmax. */
int vkrt_intersect_brute_any(vkrt_ctx *ctx, const float *tris, uint32_t n_tri, const vkrt_ray *rays,
                             uint32_t n_ray, uint8_t *occluded, int flags);

/* ---------------------------------------------------------------------------------------------
 * bounding volume hierarchy (mesh -> acceleration structure) and traversal
 * ------------------------------------------------------------------------------------------- */

enum
{
    VKRT_BUILDER_LBVH = 0, /* Morton codes + radix sort + Karras hierarchy + bottom-up refit */
    VKRT_BUILDER_PLOC = 1  /* LBVH order + parallel locally-ordered agglomerative clustering */
};

typedef struct vkrt_bvh_params
{
    uint32_t builder;       /* VKRT_BUILDER_* */
    uint32_t max_leaf_size; /* triangles per leaf, 1..8 (0 = default) */
    uint32_t reserved[6];
} vkrt_bvh_params;

typedef struct vkrt_bvh_info
{
    uint32_t n_tri;
    uint32_t n_nodes;
    uint32_t n_leaves;
    uint32_t depth;
    float sah_cost;
    float build_ms;
    uint64_t device_bytes;
    float bounds[6];
} vkrt_bvh_info;

/* Build over an indexed triangle mesh: positions with a byte stride (>= 12) between vertices,
 * 3 uint32 indices per triangle. The primitive index reported by traversal is the triangle's
 * position in `indices` (triangle j = indices[3j..3j+2]). */
int vkrt_bvh_build(vkrt_ctx *ctx, const float *positions, uint32_t n_vert, uint32_t stride_bytes,
                   const uint32_t *indices, uint32_t n_tri, const vkrt_bvh_params *params, int flags,
                   vkrt_bvh **out);
/* Rebuild bounds for moved vertices, same topology (RayBuilder's update path). */
int vkrt_bvh_refit(vkrt_ctx *ctx, vkrt_bvh *bvh, const float *positions, uint32_t stride_bytes, int flags);
int vkrt_bvh_get_info(const vkrt_bvh *bvh, vkrt_bvh_info *out);
void vkrt_bvh_destroy(vkrt_bvh *bvh);

/* Launch shape of the traversal kernels: threads per block (64, 128, 256) and traversal-stack entries
 * kept in LDS per ray (0, 8, 16, 32; deeper entries live in scratch memory). Results do not depend on
 * it. No reference counterpart: the Vulkan driver owns the traversal there. */
typedef struct vkrt_trace_config
{
    uint32_t block_size;
    uint32_t lds_stack_depth;
    uint32_t reserved[2];
} vkrt_trace_config;

int vkrt_set_trace_config(vkrt_ctx *ctx, const vkrt_trace_config *cfg);
int vkrt_get_trace_config(const vkrt_ctx *ctx, vkrt_trace_config *cfg);

/* Closest hit through the BVH; results identical to vkrt_intersect_brute_closest. */
int vkrt_trace_closest(vkrt_ctx *ctx, const vkrt_bvh *bvh, const vkrt_ray *rays, uint32_t n, vkrt_hit *out,
                       int flags);
/* Any hit (shadow query) through the BVH; results identical to vkrt_intersect_brute_any. */
int vkrt_trace_any(vkrt_ctx *ctx, const vkrt_bvh *bvh, const vkrt_ray *rays, uint32_t n, uint8_t *occluded,
                   int flags);

/* ---------------------------------------------------------------------------------------------
 * procedural geometry (host) used for the benchmark scenes
 * ------------------------------------------------------------------------------------------- */

/* Counts for Geometry::IcoSphere(radius, subdiv): vertex copies and indices. */
int vkrt_icosphere_size(uint32_t subdiv, uint32_t *n_vert, uint32_t *n_idx);
/* Fill pos (3 * n_vert floats) and idx (n_idx uint32) exactly as Geometry::IcoSphere. */
int vkrt_icosphere(float radius, uint32_t subdiv, float *pos, uint32_t *idx);
/* Geometry::Plane(width, height, segments_w, segments_h, xy_plane = false) positions and indices. */
int vkrt_plane_size(uint32_t segments_w, uint32_t segments_h, uint32_t *n_vert, uint32_t *n_idx);
int vkrt_plane(float width, float height, uint32_t segments_w, uint32_t segments_h, float *pos, uint32_t *idx);

#ifdef __cplusplus
}
#endif

#endif /* VKRT_H */
