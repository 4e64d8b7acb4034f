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
 *   vkrt_tri_box_overlap           <- intersect(const Triangle&, const AABB&)     (src/intersection.cpp:364-370)
 *   vkrt_tri_tri_overlap           <- intersect(const Triangle&, const Triangle&) (src/intersection.cpp:374-378)
 *   vkrt_compute_aabb              <- compute_aabb(const std::vector<vec3>&)      (src/intersection.cpp:96-106)
 *   vkrt_compute_bounding_sphere   <- compute_bounding_sphere(const std::vector<vec3>&) (src/intersection.cpp:25-92)
 *   vkrt_intersect_brute_*         <- a closest/any-hit loop over intersect(Triangle, Ray); the reference
 *                                     has no such loop, its callers test one triangle at a time
 *   vkrt_bvh_build / _refit        <- RayBuilder::create_mesh_structures / build_scene_acceleration
 *                                     (src/RayBuilder.cpp:98-303, :783-1035)
 *   vkrt_trace_closest / _any      <- TraceRay closest-hit / ACCEPT_FIRST_HIT shadow queries
 *                                     (shaders/slang/raypipeline.slang:182, ray/visibility_test.slang:45-87)
 *   vkrt_icosphere / vkrt_plane    <- Geometry::IcoSphere / Geometry::Plane       (src/Geometry.cpp:566-613, :313-371)
 *   vkrt_pt_*                      <- PBRPathTracer: settings_t (include/vierkant/PBRPathTracer.hpp:42-110),
 *                                     render_scene / reset_accumulator / current_batch
 *                                     (src/PBRPathTracer.cpp:184-255, :677-682), the raygen / closest-hit / miss
 *                                     programs (shaders/slang/raypipeline.slang:91-270, :360-680, :746-761) and
 *                                     pixel_data_t (shaders/slang/ray/pixel_data.slang:9-21)
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
    VKRT_ASYNC = 2,
    VKRT_CULL_BACK = 8 /* transmittance rays: skip back faces of one-sided materials */
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
/* Destroy batches and path tracers made on a context before the context: they use its stream. */
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
/* triangle (9 floats) vs AABB (6 floats: min.xyz, max.xyz) overlap; out[i] in {0, 1} */
int vkrt_tri_box_overlap(vkrt_ctx *ctx, const float *tris, const float *boxes, uint32_t n, uint32_t *out,
                         int flags);
/* triangle vs triangle overlap (boundaries count); out[i] in {0, 1} */
int vkrt_tri_tri_overlap(vkrt_ctx *ctx, const float *tris_a, const float *tris_b, uint32_t n, uint32_t *out,
                         int flags);
/* AABB of n points (3 floats each); out: 6 floats (min.xyz, max.xyz). n == 0 gives the
 * reference's empty box (min = FLT_MAX, max = -FLT_MAX). */
int vkrt_compute_aabb(vkrt_ctx *ctx, const float *points, uint32_t n, float *out6, int flags);
/* compute_bounding_sphere (src/intersection.cpp:25-92): Ritter's sphere with the reference's extremum
 * choice and growth order, bit-identical. out4 = center xyz, radius. n == 0 gives zeros. */
int vkrt_compute_bounding_sphere(vkrt_ctx *ctx, const float *points, uint32_t n, float *out4, int flags);

/* ---------------------------------------------------------------------------------------------
 * brute force: every ray against every triangle
 * ------------------------------------------------------------------------------------------- */

/* Closest hit: smallest t with tmin < t < tmax among INTERSECT results; equal t resolves to the
 * lowest triangle index. */
int vkrt_intersect_brute_closest(vkrt_ctx *ctx, const float *tris, uint32_t n_tri, const vkrt_ray *rays,
                                 uint32_t n_ray, vkrt_hit *out, int flags);
/* Allocation-free closest hit. No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Results are identical to vkrt_intersect_brute_closest. Arrays must be device pointers
 * (VKRT_DEVICE_PTRS is required). The call makes no hipMalloc/hipFree: its scratch (the triangles'
 * (v0, e1, e2) and one 64-bit key per ray, 15 MiB) is a file-scope device array that each GPU the
 * library runs on holds until the process ends. Sizes above VKRT_NOALLOC_MAX_TRI triangles or
 * VKRT_NOALLOC_MAX_RAY rays are rejected with VKRT_E_INVALID_ARG. Every call on one GPU shares that
 * GPU's array, so calls that overlap on one GPU (any contexts or streams) race. */
#define VKRT_NOALLOC_MAX_TRI 262144u
#define VKRT_NOALLOC_MAX_RAY 786432u
int vkrt_intersect_brute_closest_noalloc(vkrt_ctx *ctx, const float *tris, uint32_t n_tri, const vkrt_ray *rays,
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
    VKRT_BUILDER_PLOC = 1  /* the Morton order, then parallel locally-ordered clustering (Meister & Bittner
                            * 2018, radius 16): about 20 % lower SAH cost, 10-35 % faster traversal, builds
                            * about 25 % slower; leaves are reordered depth-first */
};

typedef struct vkrt_bvh_params
{
    uint32_t builder;       /* VKRT_BUILDER_* */
    uint32_t max_leaf_size; /* triangles per leaf, 1..8 (0 = default: 2, fastest with the 4-wide layout) */
    uint32_t width;         /* traversal layout: 0 = 4-wide collapse when its stack bound fits (default),
                             * 2 = the binary tree only. Results do not depend on it. */
    uint32_t ploc_radius;   /* PLOC neighbour search radius in sorted positions, 1..32 (0 = default: 16) */
    uint32_t reserved[4];
} vkrt_bvh_params;

/* depth, n_leaves and sah_cost are computed when the BVH is built; a refit keeps the topology and updates
   bounds (and build_ms stays the build's time) */
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
    uint32_t n_wide_nodes;  /* nodes of the 4-wide collapse (0 when traversal uses the binary tree) */
    uint32_t wide_stack;    /* worst-case traversal stack depth of the 4-wide collapse */
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

/* BVH invariants, checked on the host from the device arrays. A diagnostic with no counterpart in the
 * reference (whose acceleration structure is the driver's, RayBuilder.cpp:274): the walk traversal uses
 * (4-wide collapse or binary tree) reaches every triangle exactly once, every child box lies inside the
 * box its parent stores for it, every leaf box holds its triangles' vertices (read from `positions`, the
 * mesh the BVH was built or last refitted from), and the worst-case traversal stack recomputed from the
 * nodes equals the build's record and fits the 64-entry stack. */
typedef struct vkrt_bvh_check
{
    uint32_t ok;              /* 1 when every count below is clean */
    uint32_t wide;            /* the walk checked: 1 the 4-wide collapse, 0 the binary tree */
    uint32_t tris_reached;    /* leaf triangle records reached from the root */
    uint32_t tris_not_once;   /* triangles reached zero times or more than once */
    uint32_t box_violations;  /* child boxes outside their parent's + leaf boxes missing a vertex */
    uint32_t bad_refs;        /* child refs or leaf ranges out of bounds */
    uint32_t stack_bound;     /* worst-case stack depth recomputed from the nodes */
    uint32_t stack_ok;        /* stack_bound <= 64 and equal to the build's record */
    /* the first box violation found: kind 1 a child box outside its parent's, kind 2 a triangle vertex
     * outside its leaf's box; node and slot of the box, the triangle (kind 2), the box and the outer box or
     * vertex (kind 2: the vertex in outer[0..2]) */
    uint32_t bad_kind, bad_node, bad_slot, bad_tri;
    float bad_box[6], bad_outer[6];
} vkrt_bvh_check;
int vkrt_bvh_validate(vkrt_ctx *ctx, const vkrt_bvh *bvh, const float *positions, uint32_t stride_bytes, int flags,
                      vkrt_bvh_check *out);

/* ---------------------------------------------------------------------------------------------
 * two-level scenes: instances of mesh BVHs with per-instance transforms (the top-level structure of
 * RayBuilder::build_scene_acceleration, src/RayBuilder.cpp:386-781; entry_t::transform / inv_transform,
 * include/vierkant/RayBuilder.hpp:24-38)
 * ------------------------------------------------------------------------------------------- */

typedef struct vkrt_scene vkrt_scene;

typedef struct
{
    float object_to_world[12]; /* 3x4 row-major: world = M (p, 1) */
    float world_to_object[12]; /* its inverse (entry_t::inv_transform); traversal transforms rays with it */
    uint32_t blas;             /* index into the scene's mesh list */
    uint32_t custom_index;     /* reported with the hit */
    uint32_t reserved[2];
} vkrt_instance;

/* Closest hit of a scene (32 B): t is shared by world and object space (rays are transformed without
 * renormalizing); prim indexes the instance's mesh; instance = VKRT_INVALID_PRIM on a miss. */
typedef struct
{
    float t, u, v;
    uint32_t prim;
    uint32_t instance;
    uint32_t custom_index;
    uint32_t reserved[2];
} vkrt_instance_hit;

/* Fill object_to_world and its inverse (computed in double, rounded once) of an instance. */
void vkrt_instance_set_transform(vkrt_instance *inst, const float object_to_world[12]);
/* Build the top level over `instances` of `meshes` (4-wide mesh BVHs). Ties: smallest t, then the lowest
 * instance, then the lowest triangle. The meshes must outlive the scene; after a mesh refit or a
 * transform change, call vkrt_scene_set_instances (the top level is rebuilt on the host). */
int vkrt_scene_create(vkrt_ctx *ctx, vkrt_bvh *const *meshes, uint32_t n_meshes, const vkrt_instance *instances,
                      uint32_t n_instances, vkrt_scene **out);
int vkrt_scene_set_instances(vkrt_scene *scene, const vkrt_instance *instances, uint32_t n_instances);
/* Before the context and its meshes. */
void vkrt_scene_destroy(vkrt_scene *scene);
int vkrt_scene_trace_closest(vkrt_ctx *ctx, vkrt_scene *scene, const vkrt_ray *rays, uint32_t n,
                             vkrt_instance_hit *out, int flags);
int vkrt_scene_trace_any(vkrt_ctx *ctx, vkrt_scene *scene, const vkrt_ray *rays, uint32_t n, uint8_t *occluded,
                         int flags);
/* instance count, top-level nodes, worst-case traversal stack (top level + deepest mesh, at most 96) */
int vkrt_scene_get_info(const vkrt_scene *scene, uint32_t *n_instances, uint32_t *n_nodes, uint32_t *stack_bound);

/* Launch shape of the traversal kernels: threads per block (64, 128, 256), traversal-stack entries
 * kept in LDS per ray (0, 8, 16, 32; deeper entries live in scratch memory), the loop of the 4-wide
 * walk (VKRT_TRAVERSAL_*) and the order in which workgroups take rays: xcd_chunk 0 gives workgroup b the
 * rays [b B, (b + 1) B); xcd_chunk C (a power of two up to 256) hands C consecutive ray blocks to the
 * workgroups that the round-robin dispatch sends to one XCD, so neighbouring rays share that XCD's L2.
 * Results do not depend on it. No reference counterpart: the Vulkan driver owns the traversal there. */
typedef struct vkrt_trace_config
{
    uint32_t block_size;
    uint32_t lds_stack_depth;
    uint32_t traversal;
    uint32_t xcd_chunk;
} vkrt_trace_config;

enum
{
    VKRT_TRAVERSAL_STEP = 0,     /* one node or one leaf per loop step */
    VKRT_TRAVERSAL_POSTPONED = 1 /* leaves parked until every lane of the wave holds one, then tested together */
};


int vkrt_set_trace_config(vkrt_ctx *ctx, const vkrt_trace_config *cfg);
int vkrt_get_trace_config(const vkrt_ctx *ctx, vkrt_trace_config *cfg);

/* Closest hit through the BVH; results identical to vkrt_intersect_brute_closest. */
int vkrt_trace_closest(vkrt_ctx *ctx, const vkrt_bvh *bvh, const vkrt_ray *rays, uint32_t n, vkrt_hit *out,
                       int flags);
/* Any hit (shadow query) through the BVH; results identical to vkrt_intersect_brute_any. */
int vkrt_trace_any(vkrt_ctx *ctx, const vkrt_bvh *bvh, const vkrt_ray *rays, uint32_t n, uint8_t *occluded,
                   int flags);

/* ---------------------------------------------------------------------------------------------
 * batched closest hit over many scenes: one persistent launch, replayed as a captured hipGraph
 * (no reference counterpart: vierkant traces one scene per vkCmdTraceRaysKHR, src/RayTracer.cpp:128-129)
 * ------------------------------------------------------------------------------------------- */

typedef struct vkrt_batch vkrt_batch;

/* one scene's work: its BVH, and device buffers of n_rays rays in and hit records out */
typedef struct vkrt_batch_item
{
    const vkrt_bvh *bvh;
    const vkrt_ray *rays;
    vkrt_hit *hits;
    uint32_t n_rays;
    uint32_t reserved;
} vkrt_batch_item;

/* launch shape: threads per block (64/128/256), traversal-stack entries in LDS (0/8/16/32), rays per tile
 * (a multiple of 64) and persistent blocks per compute unit (1..64) */
typedef struct vkrt_batch_config
{
    uint32_t block_size;      /* 64, 128 or 256 */
    uint32_t lds_stack_depth; /* top entries of the 64-entry traversal stack kept in LDS: 0, 8, 16, 24, 32, 48, 64 */
    uint32_t tile;            /* rays per claimed work item, a multiple of 64 */
    uint32_t blocks_per_cu;   /* persistent blocks launched per compute unit */
    uint32_t xcd_chunk;       /* 0: one queue in tile order; C: chunks of C tiles dealt to the 8 XCDs' queues */
    uint32_t reserved[3];
} vkrt_batch_config;

/* Capture the batch (items are copied; their buffers must stay valid) into a graph with the default shape. */
int vkrt_batch_create(vkrt_ctx *ctx, const vkrt_batch_item *items, uint32_t n_items, vkrt_batch **out);
/* Before the context and before the BVHs and buffers its items name. */
void vkrt_batch_destroy(vkrt_batch *batch);
/* Change the launch shape (re-captures the graph). Results do not depend on it. */
int vkrt_batch_set_config(vkrt_batch *batch, const vkrt_batch_config *cfg);
int vkrt_batch_get_config(const vkrt_batch *batch, vkrt_batch_config *out);
/* Replay the graph: closest hits of every item's rays into its hit buffer. */
int vkrt_batch_launch(vkrt_batch *batch, int flags);
/* Coordinate descent over the launch shapes of vkrt_batch_candidates (block size, LDS stack depth 16-64, tile,
 * threads per CU, XCD chunk) from the batch's current shape, `repeats` replays per shape; keeps the fastest
 * shape found and returns it and its time per launch. */
int vkrt_batch_autotune(vkrt_batch *batch, uint32_t repeats, vkrt_batch_config *best, float *best_ms);
/* The autotuner's shape space: writes up to `capacity` shapes, returns how many there are. The autotuner
 * (a coordinate descent from the batch's current shape) only ever picks one of these. */
uint32_t vkrt_batch_candidates(vkrt_batch_config *out, uint32_t capacity);
uint64_t vkrt_batch_total_rays(const vkrt_batch *batch);
/* One-shot: create, launch once, destroy. */
int vkrt_batch_closest(vkrt_ctx *ctx, const vkrt_batch_item *items, uint32_t n_items, int flags);

/* ---------------------------------------------------------------------------------------------
 * path tracer (PBRPathTracer)
 * ------------------------------------------------------------------------------------------- */

typedef struct vkrt_pt vkrt_pt;

/* PBRPathTracer::settings_t fields that affect the traced image (PBRPathTracer.hpp:42-110). Defaults
 * (vkrt_pt_default_settings): 1280x720, max_num_batches 0, num_samples 1, max_trace_depth 6,
 * max_path_beta 0, suspend_trace_when_done 1, disable_material 0, suppress_reset 0, denoising 0, tonemap 1,
 * gamma 1, exposure 2, mis_mode 0, suppress_refractive_caustics 0 (include/vierkant/PBRPathTracer.hpp:42-90). Bloom is not provided: the composited
 * image is the reference's with settings_t::bloom false. */
typedef struct vkrt_pt_settings
{
    uint32_t resolution[2];
    uint32_t max_num_batches;
    uint32_t num_samples;
    uint32_t max_trace_depth;
    float max_path_beta;
    uint32_t suspend_trace_when_done;
    uint32_t disable_material;
    uint32_t suppress_reset;
    uint32_t denoising;     /* settings_t::denoising: vkrt_pt_read_image applies the a-trous denoiser */
    uint32_t tonemap;       /* settings_t::tonemap: vkrt_pt_read_image composites as bloom_composition.slang */
    float gamma;            /* settings_t::gamma (> 0 with tonemap) */
    float exposure;         /* settings_t::exposure */
    uint32_t mis_mode;      /* settings_t::mis_mode: VKRT_MIS_FULL, _NEE_ONLY, _BSDF_ONLY (ray_common.slang:289-294) */
    uint32_t suppress_refractive_caustics; /* settings_t::suppress_refractive_caustics */
    uint32_t reserved[1];
} vkrt_pt_settings;

/* Direct-light estimator (MisMode, shaders/slang/ray/ray_common.slang:289-294). */
enum
{
    VKRT_MIS_FULL = 0,      /* next-event estimation and BSDF sampling, power heuristic */
    VKRT_MIS_NEE_ONLY = 1,  /* light bodies reached by BSDF rays add nothing */
    VKRT_MIS_BSDF_ONLY = 2  /* next-event samples add nothing for lights BSDF rays can reach */
};

/* LightType (include/vierkant/punctual_light.hpp:20-33, shaders/slang/ray/ray_common.slang:92-107). */
enum
{
    VKRT_LIGHT_OMNI = 0,
    VKRT_LIGHT_SPOT = 1,
    VKRT_LIGHT_DIRECTIONAL = 2, /* angular_size > 0: a sun disc of that apex angle */
    VKRT_LIGHT_AREA = 3,        /* reserved by the reference for emissive triangles; never gathered */
    VKRT_LIGHT_RECT = 4,
    VKRT_LIGHT_SPHERE = 5,
    VKRT_LIGHT_TUBE = 6,
    VKRT_LIGHT_DISK = 7
};

/* light_t, the packed light the shaders read (include/vierkant/punctual_light.hpp:64-89; 80 bytes).
 * direction is the propagation direction; color * intensity is the emitted radiance of light bodies and the
 * intensity of punctual lights; size_x is the radius (sphere, disk, tube) or half-width (rect), size_y the
 * half-height (rect) or half-length (tube); tangent is the rect's in-plane x axis. */
typedef struct vkrt_light
{
    float position[3];
    uint32_t type;
    float color[3];
    float intensity;
    float direction[3];
    float range;
    float spot_angle_scale, spot_angle_offset;
    float angular_size;
    float size_y;
    float tangent[3];
    float size_x;
} vkrt_light;

/* material_t subset evaluated by the tracer (shaders/slang/ray/types.slang:64-107): base color, emission
 * (rgb, a = strength), metal/roughness, transmission + ior, clearcoat, sheen, diffuse transmission. */
typedef struct vkrt_material
{
    float color[4];
    float emission[4];
    float metalness, roughness, transmission, ior;
    float clearcoat, clearcoat_roughness, sheen_roughness, diffuse_transmission;
    float sheen_color[4];
    float diffuse_transmission_color[4];
    uint32_t two_sided; /* its triangles are never culled as back faces (VK_GEOMETRY_INSTANCE_TRIANGLE_FACING_CULL_DISABLE) */
    uint32_t color_texture; /* 1 + index into the scene's textures (0: none): color and alpha are multiplied by
                             * the texture, sampled trilinearly with repeat addressing at the ray-cone level of
                             * detail (raypipeline.slang:300-321, :526-531) */
    uint32_t alpha_mode;    /* VKRT_ALPHA_*: the any-hit test of every ray (MASK: alpha < alpha_cutoff skips the
                             * surface; BLEND: alpha < a uniform draw skips it) */
    float alpha_cutoff;
    /* further texture slots (1 + texture index, 0: none), sampled at the same level of detail as the color
     * (raypipeline.slang:536-606, material_t texture slots of ray/types.slang:85-103) */
    uint32_t normal_texture;        /* tangent-space normal in rg (z = sqrt(1 - x^2 - y^2)); needs tangents */
    uint32_t emission_texture;      /* rgb multiplies the emission, or replaces it where the emission is black */
    uint32_t ao_rough_metal_texture; /* g scales roughness, b metalness (glTF metallicRoughnessTexture) */
    uint32_t transmission_texture;  /* r scales transmission */
    uint32_t diffuse_transmission_texture;       /* a scales diffuse_transmission */
    uint32_t diffuse_transmission_color_texture; /* rgb tints diffuse_transmission_color */
    /* KHR_materials_volume (include/vierkant/Material.hpp:104-108): an interior medium absorbing as
     * exp(-sigma_t d) per channel, sigma_t = -log(attenuation_color) / attenuation_distance
     * (raypipeline.slang:380-415, :493, visibility_test.slang:110-129). attenuation_distance 0 stands for the
     * reference's default +inf: no interior medium */
    float attenuation_color[3];
    float attenuation_distance;
    uint32_t reserved[2];
} vkrt_material;

/* BlendMode (include/vierkant/Material.hpp:129): the any-hit stages of raypipeline.slang:686-720 and
 * visibility_test.slang:136-155. */
enum
{
    VKRT_ALPHA_OPAQUE = 0,
    VKRT_ALPHA_MASK = 1,
    VKRT_ALPHA_BLEND = 2
};

/* An RGBA8 UNORM image, rows from the top, and its sampler (texture_sampler_t, include/vierkant/Material.hpp:
 * 157-181; glTF samplers map to it as src/model/gltf.cpp:662-690 does). The tracer builds the mip chain (2x2
 * box filter, as the reference's linear blits of a level of even size) and blends levels linearly. Zero
 * fields give the defaults: repeat addressing, linear filtering. */
typedef struct vkrt_texture
{
    uint32_t width, height;
    const uint8_t *rgba;
    uint32_t address_mode_u, address_mode_v; /* VKRT_ADDRESS_* */
    uint32_t min_filter, mag_filter;         /* VKRT_FILTER_*: min applies where lambda > 0, mag elsewhere */
} vkrt_texture;

enum
{
    VKRT_ADDRESS_REPEAT = 0,
    VKRT_ADDRESS_MIRRORED_REPEAT = 1,
    VKRT_ADDRESS_CLAMP_TO_EDGE = 2
};

enum
{
    VKRT_FILTER_LINEAR = 0,
    VKRT_FILTER_NEAREST = 1
};

/* A flat scene: an indexed triangle mesh with shading normals, split into objects (entities) that each
 * carry one material. All pointers are host pointers; the tracer keeps device copies. */
typedef struct vkrt_pt_scene
{
    const float *positions; /* 3 floats per vertex */
    const float *normals;   /* 3 floats per vertex */
    uint32_t n_vert;
    const uint32_t *indices; /* 3 per triangle */
    uint32_t n_tri;
    const uint32_t *tri_object;      /* object index per triangle */
    const uint32_t *object_material; /* material index per object */
    uint32_t n_objects;
    const vkrt_material *materials;
    uint32_t n_materials;
    float environment; /* scale of the procedural sky (environment_white) */
    const float *texcoords;        /* 2 per vertex, or null (all 0) */
    const vkrt_texture *textures;  /* referenced by the material texture slots */
    uint32_t n_textures;
    const float *tangents;         /* 3 per vertex (normal maps), or null (all 0) */
} vkrt_pt_scene;

/* Scene files into a flat scene for the tracer: Wavefront OBJ + MTL and glTF 2.0 (.gltf with external or
 * base64 buffers, .glb), the formats of vierkant::model::wavefront_obj and vierkant::model::gltf
 * (src/model/wavefront_obj.cpp:11-137, src/model/gltf.cpp:136-660, :933-1218). OBJ: every face corner is a
 * vertex, polygons are fanned, missing normals are computed as Geometry::compute_vertex_normals, one object
 * per shape with the material of its first face. glTF: one object per triangle primitive of every mesh node
 * reachable from the scene, node transforms baked into world-space vertices (double precision, rounded
 * once), material factors of the core spec and of KHR_materials_emissive_strength / transmission / ior /
 * clearcoat / sheen / diffuse_transmission, PNG textures in every slot the tracer samples (base color with
 * alphaMode / alphaCutoff, metallicRoughness, normal, emissive, transmission, diffuse transmission and its
 * color) with the base color texture's KHR_texture_transform folded into the coordinates, tangents from
 * TANGENT or Geometry::compute_tangents. OBJ: map_Kd and norm textures; a map_Kd material's coordinates are
 * flipped vertically as the reference's texture transform does (wavefront_obj.cpp:100-111). Host code only:
 * no context or GPU is needed. Errors: VKRT_E_INVALID_ARG with the message in vkrt_model_error (per thread). */
typedef struct vkrt_model vkrt_model;
typedef struct vkrt_model_view
{
    const float *positions; /* 3 per vertex, world space */
    const float *normals;   /* 3 per vertex */
    const float *texcoords; /* 2 per vertex (0 where the file has none) */
    const float *tangents;  /* 3 per vertex: the file's (glTF TANGENT xyz) or Geometry::compute_tangents */
    uint32_t n_vert;
    const uint32_t *indices; /* 3 per triangle */
    uint32_t n_tri;
    const uint32_t *tri_object;
    const uint32_t *object_material;
    uint32_t n_objects;
    const vkrt_material *materials;
    uint32_t n_materials;
    const vkrt_texture *textures; /* decoded PNG images (any bit depth and color type, Adam7 included) */
    uint32_t n_textures;
    uint32_t n_skipped_textures;  /* images that are not PNG or failed to decode; their materials are untextured */
} vkrt_model_view;
int vkrt_model_load(const char *path, vkrt_model **out);
const char *vkrt_model_error(void);
/* Views stay valid until vkrt_model_destroy. */
int vkrt_model_get(const vkrt_model *model, vkrt_model_view *out);
/* The view as the tracer's scene (vkrt_pt_set_scene), with the given sky scale. */
int vkrt_model_pt_scene(const vkrt_model *model, float environment, vkrt_pt_scene *out);
void vkrt_model_destroy(vkrt_model *model);

/* Pinhole / thin-lens camera (camera_params_t): eye, unit forward and up vectors, vertical field of view
 * in radians, aperture diameter (0 = pinhole) and focal distance. */
typedef struct vkrt_camera
{
    float eye[3];
    float fov_y;
    float forward[3];
    float aperture;
    float up[3];
    float focal_distance;
} vkrt_camera;

/* pixel_data_t (32 B): accumulated radiance, this batch's coverage, projected depth, octahedral normal,
 * object id (0xFFFF on a miss), albedo RGBA8 of the first hit. */
typedef struct vkrt_pixel
{
    float radiance[3];
    float alpha;
    float depth;
    uint32_t normal;
    uint32_t object_id;
    uint32_t albedo;
} vkrt_pixel;

typedef struct vkrt_pt_stats
{
    float raytrace_ms;       /* GPU time of the last traced batch */
    uint32_t traced;         /* 1 if the last vkrt_pt_render traced a batch */
    uint64_t segments;       /* path segments (extension rays) traced in the last batch */
    uint64_t samples;        /* camera samples in the last batch */
} vkrt_pt_stats;

void vkrt_pt_default_settings(vkrt_pt_settings *out);
/* seed: create_info_t::seed; each traced batch draws its random seed from std::default_random_engine(seed) */
int vkrt_pt_create(vkrt_ctx *ctx, const vkrt_pt_settings *settings, uint32_t seed, vkrt_pt **out);
/* Before the context. */
void vkrt_pt_destroy(vkrt_pt *pt);
int vkrt_pt_set_settings(vkrt_pt *pt, const vkrt_pt_settings *settings);
int vkrt_pt_get_settings(const vkrt_pt *pt, vkrt_pt_settings *out);
/* Upload the scene and build its BVH (the RayBuilder step of the reference's render_scene). */
int vkrt_pt_set_scene(vkrt_pt *pt, const vkrt_pt_scene *scene);

/* An instance of a mesh of an instanced scene: entry_t's transform (include/vierkant/RayBuilder.hpp:24-38,
 * shaders/slang/ray/types.slang:17-32; a vierkant::transform_t, include/vierkant/transform.hpp): points map
 * to rotate(rotation, scale * p) + translation, normals and tangents turn by the rotation
 * (raypipeline.slang:336-358). rotation is a unit quaternion (x, y, z, w). */
typedef struct vkrt_pt_instance
{
    float translation[3];
    float rotation[4];
    float scale[3];
    uint32_t mesh; /* index into the scene's meshes */
    uint32_t reserved[3];
} vkrt_pt_instance;

/* An instanced scene (the two-level build of RayBuilder, src/RayBuilder.cpp:386-781, and the per-entry
 * transforms PBRPathTracer hands it, src/PBRPathTracer.cpp:655-675): `scene` holds every mesh's geometry in its
 * own object space, mesh m being triangles [mesh_first[m], mesh_first[m + 1]) (n_meshes + 1 offsets, from 0 to
 * scene->n_tri); each mesh gets its own BVH, and `instances` place meshes in the world (a mesh may have any
 * number of instances, or none). Materials, objects and textures are the scene's. The instances are walked
 * through a top-level BVH; a hit's instance transforms its shading position, normal and tangent, its
 * triangle area for texture level of detail, and the scale of the object box for the in-volume tmax
 * (raypipeline.slang:370). Ties go to the smallest t, then the lowest instance, then the lowest triangle.
 * Replaces the scene and BVH of vkrt_pt_set_scene (and vice versa). */
int vkrt_pt_set_instanced_scene(vkrt_pt *pt, const vkrt_pt_scene *scene, const uint32_t *mesh_first,
                                uint32_t n_meshes, const vkrt_pt_instance *instances, uint32_t n_instances);
/* The frame's light list (PBRPathTracer::update_trace_descriptors, src/PBRPathTracer.cpp:607-641): lights of
 * positive intensity, other than VKRT_LIGHT_AREA, are kept in the given order (a sun first, as the reference
 * folds it in). With n > 0 after that filter every surface vertex adds next-event estimation with a
 * transmittance shadow ray (ray/direct_lighting.slang:32-366), and light bodies met by BSDF rays add their
 * MIS-weighted radiance (raypipeline.slang:201-222, :728-761). At most 65536 lights; n = 0 clears the list. */
int vkrt_pt_set_lights(vkrt_pt *pt, const vkrt_light *lights, uint32_t n);
/* Launch shape of the path tracer's kernels (no counterpart in the reference, whose RT pipeline schedules
 * itself). mode FULL_GRID (the default): one workgroup per tile of the pixel count at every depth, the tiles
 * past the live queue returning at once; PERSISTENT: the per-depth trace and shadow kernels are fixed grids of
 * the workgroups the GPU holds at once that claim the live queue's tiles with atomics. tail_depth (>= 1;
 * default VKRT_PT_TAIL_DEPTH_DEFAULT): outside the radiance-cache mode the paths alive at that depth finish in
 * one lane each, every remaining depth in one kernel; VKRT_PT_NO_TAIL keeps one launch per depth to the end.
 * Every combination gives the same image bit for bit. */
enum
{
    VKRT_PT_LAUNCH_PERSISTENT = 0,
    VKRT_PT_LAUNCH_FULL_GRID = 1
};
#define VKRT_PT_TAIL_DEPTH_DEFAULT 4u
#define VKRT_PT_NO_TAIL 0xffffffffu
int vkrt_pt_set_launch_mode(vkrt_pt *pt, int mode, uint32_t tail_depth);
/* One render_scene call: traces one batch of settings.num_samples samples per pixel into the running mean
 * unless max_num_batches is reached with suspend_trace_when_done set; then advances the batch index as
 * min(batch + 1, max_num_batches). A resolution change resets the accumulator. */
int vkrt_pt_render(vkrt_pt *pt, const vkrt_camera *camera);
/* Reset the batch index to 0 unless settings.suppress_reset is set (PBRPathTracer.cpp:677-680). */
int vkrt_pt_reset_accumulator(vkrt_pt *pt);
uint32_t vkrt_pt_current_batch(const vkrt_pt *pt);
/* Copy the W*H pixel buffer (vkrt_pixel, row-major from the top-left) to `out` (host, or device with flags). */
int vkrt_pt_read_pixels(vkrt_pt *pt, vkrt_pixel *out, int flags);
/* The tracer's output image (W*H float4, row-major from the top-left): with settings.denoising the
 * edge-avoiding a-trous wavelet of denoise.slang over the pixel buffer (5 passes, steps 1..16, rgb, alpha 0;
 * PBRPathTracer::denoise_pass, src/PBRPathTracer.cpp:346-400), otherwise radiance and alpha. With
 * settings.tonemap that image is then composited as PBRPathTracer::post_fx_pass (src/PBRPathTracer.cpp:405-465,
 * shaders/slang/fullscreen/bloom_composition.slang:34-45, no bloom, no motion): rgb = pow(1 - exp(-rgb *
 * exposure), 1 / gamma), alpha kept, each channel rounded to half precision as the R16G16B16A16_SFLOAT target
 * stores it (src/PBRPathTracer.cpp:819). */
int vkrt_pt_read_image(vkrt_pt *pt, float *out_rgba, int flags);
/* That composite alone over n float4 texels (in and out may alias). */
int vkrt_tonemap(vkrt_ctx *ctx, const float *in_rgba, uint32_t n, float exposure, float gamma, float *out_rgba,
                 int flags);
/* The same denoise pass over any pixel buffer (width * height vkrt_pixel). */
int vkrt_denoise(vkrt_ctx *ctx, const vkrt_pixel *pixels, uint32_t width, uint32_t height, uint32_t denoise,
                 float *out_rgba, int flags);
int vkrt_pt_get_stats(const vkrt_pt *pt, vkrt_pt_stats *out);
/* transmittance_test (shaders/slang/ray/visibility_test.slang:94-165) through the tracer's scene: out4 =
 * (rgb transmittance, 0) per ray; 0 when an opaque triangle (transmission == 0) is hit in (tmin, tmax)
 * and not cut out by its alpha mask (MASK: alpha below the cutoff lets the ray through; BLEND surfaces block,
 * as this call draws no random numbers), else the product of transmission * color over every transmissive hit, taken in ascending (t, triangle)
 * order (the reference's order is the driver's). VKRT_CULL_BACK in flags culls back faces of one-sided
 * materials (the RAY_FLAG_CULL_BACK_FACING_TRIANGLES call, direct_lighting.slang:354). Materials have no
 * interior absorption here (attenuation_distance = infinity). */
int vkrt_pt_trace_transmittance(vkrt_pt *pt, const vkrt_ray *rays, uint32_t n, float *out4, int flags);

/* ---------------------------------------------------------------------------------------------
 * radiance cache: multiresolution hash grid (16 levels x 2^log2_table entries x 2 features, fp16) +
 * MLP 64 -> 5 x 64 ReLU -> 3 (bf16 matrix inputs, fp32 accumulation), trained online with Adam on a
 * relative L2 loss. No counterpart in the reference (a new component, not a drop-in).
 *
 * Record (16 floats per sample): pos.xyz, roughness | wo.xyz (unit, towards the viewer), metalness |
 * shading normal.xyz, transmission | albedo.rgb, unused. Targets: 4 floats (rgb radiance, unused).
 * ------------------------------------------------------------------------------------------- */

typedef struct vkrt_rc vkrt_rc;

typedef struct
{
    uint32_t log2_table;      /* hash table entries per level = 2^log2_table, 10..24 (default 19) */
    uint32_t base_resolution; /* coarsest grid resolution (16) */
    uint32_t max_resolution;  /* finest grid resolution (2048) */
    float learning_rate;      /* Adam (1e-2) */
    float beta1, beta2;       /* (0.9, 0.99) */
    float epsilon;            /* (1e-15) */
    uint32_t reserved[9];
} vkrt_rc_params;

/* vkrt_rc_train flag: compute gradients only (readable with vkrt_rc_get_gradients), no Adam step */
#define VKRT_RC_GRAD_ONLY 4

void vkrt_rc_default_params(vkrt_rc_params *out);
/* Parameters: grid U(-1e-4, 1e-4), Glorot-uniform MLP, both from `seed`. Bounds default to the unit cube. */
int vkrt_rc_create(vkrt_ctx *ctx, const vkrt_rc_params *params, uint32_t seed, vkrt_rc **out);
/* Before the context. */
void vkrt_rc_destroy(vkrt_rc *rc);
/* parameters, MLP weight count (21,504) and grid parameter count */
int vkrt_rc_get_config(const vkrt_rc *rc, vkrt_rc_params *params, uint32_t *n_mlp, uint64_t *n_grid);
/* World-space box mapped to the grid's unit cube (positions outside are clamped). */
int vkrt_rc_set_bounds(vkrt_rc *rc, const float bmin[3], const float bmax[3]);
/* out[4 i .. 4 i + 2] = max(prediction, 0) for record i; out[4 i + 3] = 0. */
int vkrt_rc_query(vkrt_rc *rc, const float *records, uint32_t n, float *out, int flags);
/* One optimisation step on n (record, target) pairs; *loss (nullable) = mean relative L2 before the step. */
int vkrt_rc_train(vkrt_rc *rc, const float *records, const float *targets, uint32_t n, int flags, float *loss);
/* fp32 master parameters (host arrays of n_mlp and n_grid floats; either may be NULL). MLP layout: W1..W5
 * [64 out][64 in], W6 [16 out][64 in] (outputs 3..15 are kept at zero); grid: [level][entry][feature].
 * set_params also resets the Adam moments and step count. */
int vkrt_rc_get_params(vkrt_rc *rc, float *mlp, float *grid);
int vkrt_rc_set_params(vkrt_rc *rc, const float *mlp, const float *grid);
/* Gradients of the last vkrt_rc_train call made with VKRT_RC_GRAD_ONLY (same layouts as the parameters). */
int vkrt_rc_get_gradients(vkrt_rc *rc, float *mlp, float *grid);
/* Adam steps taken */
uint32_t vkrt_rc_step(const vkrt_rc *rc);

/* Path tracing with the radiance cache (no counterpart in the reference). */
typedef struct
{
    uint32_t cache_depth;       /* paths stop at this hit depth and add beta * cache (2: after two bounces) */
    uint32_t train_stride;      /* one full-length training path per train_stride pixels per sample (64) */
    uint32_t max_train_records; /* training records (path vertices) per sample, at most (16384) */
    uint32_t reserved[5];
} vkrt_pt_cache_config;

void vkrt_pt_default_cache_config(vkrt_pt_cache_config *out);
/* Attach `rc` (not owned; NULL detaches) to the tracer: every sample then queries the cache at
 * cache_depth and trains it once on the sample's training paths. The cache's bounds are set to the
 * scene's, padded by 1%. Results are not bitwise reproducible across runs (training records are appended
 * in arrival order). */
int vkrt_pt_set_radiance_cache(vkrt_pt *pt, vkrt_rc *rc, const vkrt_pt_cache_config *cfg);

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
