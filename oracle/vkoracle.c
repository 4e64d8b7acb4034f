/*
 * vkoracle.c - CPU oracle for the ray-intersection path (TEST INFRASTRUCTURE ONLY).
 *
 * A plain-C restatement of the reference's arithmetic, used as the comparison reference by tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py. The product never links it.
 *
 * glm semantics restated (g-truc/glm, the reference's math library; its submodule is empty in the
 * reference checkout, so the functions below follow glm's published source):
 *   dot(a, b)     = (a.x*b.x + a.y*b.y) + a.z*b.z               (compute_dot<vec3>: tmp = a*b; x+y+z)
 *   cross(x, y)   = (x.y*y.z - y.y*x.z, x.z*y.x - y.z*x.x, x.x*y.y - y.x*x.y)
 *   normalize(v)  = v * (1 / sqrt(dot(v, v)))                    (inversesqrt = 1 / sqrt)
 *   mat4 * vec4   = (m0*v.x + m1*v.y) + (m2*v.z + m3*v.w)         (column-major, per component)
 * Compiled with -ffp-contract=off so no multiply-add is fused (the reference's Release flags have no
 * -march, i.e. x86-64 SSE2 without FMA).
 */
#include "vkoracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef struct
{
    float x, y, z;
} v3;

typedef struct
{
    double x, y, z;
} v3d;

static inline v3 mk(const float *p) { return (v3){p[0], p[1], p[2]}; }
static inline v3 sub(v3 a, v3 b) { return (v3){a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline v3 add(v3 a, v3 b) { return (v3){a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline v3 scl(v3 a, float s) { return (v3){a.x * s, a.y * s, a.z * s}; }
static inline float dot(v3 a, v3 b)
{
    float mx = a.x * b.x, my = a.y * b.y, mz = a.z * b.z;
    return (mx + my) + mz;
}
static inline v3 cross(v3 x, v3 y)
{
    return (v3){x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y};
}
static inline v3 normalize(v3 v)
{
    float inv = 1.0f / sqrtf(dot(v, v));
    return scl(v, inv);
}

static inline v3d mkd(const float *p) { return (v3d){p[0], p[1], p[2]}; }
static inline v3d subd(v3d a, v3d b) { return (v3d){a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline double dotd(v3d a, v3d b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline v3d crossd(v3d x, v3d y)
{
    return (v3d){x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y};
}

/* ------------------------------------------------------------------------------------------------
 * ray - triangle: src/intersection.cpp:114-129
 * ---------------------------------------------------------------------------------------------- */
int orc_intersect_tri(const float *tri, const float *o, const float *d, float *t, float *u, float *v)
{
    v3 v0 = mk(tri), v1 = mk(tri + 3), v2 = mk(tri + 6);
    v3 dir = mk(d), org = mk(o);
    v3 e1 = sub(v1, v0), e2 = sub(v2, v0);                /* :116 */
    v3 pvec = cross(dir, e2);                             /* :117 */
    float det = dot(e1, pvec);                            /* :118 */
    const float epsilon = 10.0e-10f;                      /* :119 */
    if(det > -epsilon && det < epsilon) return 0;         /* :120 */
    float inv_det = 1.0f / det;                           /* :121 */
    v3 tvec = sub(org, v0);                               /* :122 */
    float uu = inv_det * dot(tvec, pvec);                 /* :123 */
    if(uu < 0.0f || uu > 1.0f) return 0;                  /* :124 */
    v3 qvec = cross(tvec, e1);                            /* :125 */
    float vv = dot(dir, qvec) * inv_det;                  /* :126 */
    if(vv < 0.0f || (uu + vv) > 1.0f) return 0;           /* :127 */
    *t = dot(e2, qvec) * inv_det;                         /* :128 */
    *u = uu;
    *v = vv;
    return 1;
}

int orc_intersect_tri_f64(const float *tri, const float *o, const float *d, double *t, double *u, double *v)
{
    v3d v0 = mkd(tri), v1 = mkd(tri + 3), v2 = mkd(tri + 6);
    v3d dir = mkd(d), org = mkd(o);
    v3d e1 = subd(v1, v0), e2 = subd(v2, v0);
    v3d pvec = crossd(dir, e2);
    double det = dotd(e1, pvec);
    const double epsilon = (double)10.0e-10f;
    if(det > -epsilon && det < epsilon) return 0;
    double inv_det = 1.0 / det;
    v3d tvec = subd(org, v0);
    double uu = inv_det * dotd(tvec, pvec);
    if(uu < 0.0 || uu > 1.0) return 0;
    v3d qvec = crossd(tvec, e1);
    double vv = dotd(dir, qvec) * inv_det;
    if(vv < 0.0 || (uu + vv) > 1.0) return 0;
    *t = dotd(e2, qvec) * inv_det;
    *u = uu;
    *v = vv;
    return 1;
}

/* ------------------------------------------------------------------------------------------------
 * ray - OBB slab test: src/intersection.cpp:152-184
 * ---------------------------------------------------------------------------------------------- */
int orc_intersect_obb(const float *obb, const float *o, const float *d, float *dist)
{
    const float epsilon = 1e-10f;                          /* :154 */
    float t_min = -FLT_MAX;                                /* :156 numeric_limits<float>::lowest() */
    float t_max = FLT_MAX;                                 /* :157 */
    v3 center = mk(obb), dir = mk(d);
    v3 p = sub(center, mk(o));                             /* :158 */
    for(int i = 0; i < 3; i++)
    {
        v3 axis = mk(obb + 3 + 3 * i);
        float h = obb[12 + i];
        float e = dot(axis, p);                            /* :162 */
        float f = dot(axis, dir);                          /* :163 */
        if(fabsf(f) > epsilon)                             /* :166 */
        {
            float t1 = (e + h) / f;                        /* :168 */
            float t2 = (e - h) / f;                        /* :169 */
            if(t1 > t2)
            {
                float tmp = t1;
                t1 = t2;
                t2 = tmp;
            }
            if(t1 > t_min) t_min = t1;
            if(t2 < t_max) t_max = t2;
            if(t_min > t_max) return 0;
            if(t_max < 0) return 0;
        }
        else if((-e - h) > 0 || (-e + h) < 0) { return 0; } /* :177 */
    }
    *dist = (t_min > 0) ? t_min : t_max;                   /* :179-183 */
    return 1;
}

/* glm::mat4(1) * vec4(c, 1): each component is (m0*x + m1*y) + (m2*z + m3*1) with 0/1 entries */
static inline float ident_row(const float c[3], int row)
{
    float m[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    float x = m[0][row] * c[0], y = m[1][row] * c[1], z = m[2][row] * c[2], w = m[3][row] * 1.0f;
    return (x + y) + (z + w);
}

/* AABB::intersect(const Ray&): src/intersection.cpp:308-312 -> OBB(aabb, mat4(1)) :219-227 */
int orc_intersect_aabb(const float *box6, const float *o, const float *d, float *dist)
{
    v3 mn = mk(box6), mx = mk(box6 + 3);
    /* AABB::center() = (max + min) / 2.f ; half_extents() = (max - min) / 2.f (intersection.hpp:297,301) */
    float c[3] = {(mx.x + mn.x) / 2.f, (mx.y + mn.y) / 2.f, (mx.z + mn.z) / 2.f};
    float he[3] = {(mx.x - mn.x) / 2.f, (mx.y - mn.y) / 2.f, (mx.z - mn.z) / 2.f};
    float obb[15];
    for(int r = 0; r < 3; ++r) obb[r] = ident_row(c, r);
    /* axis[i] = normalize(t[i].xyz()) of the identity: unit axes; scale = length(t[i]) = 1 */
    const float axes[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for(int i = 0; i < 9; ++i)
    {
        int a = i / 3;
        v3 col = {axes[3 * a], axes[3 * a + 1], axes[3 * a + 2]};
        v3 n = normalize(col);
        obb[3 + i] = (i % 3 == 0) ? n.x : (i % 3 == 1) ? n.y : n.z;
    }
    for(int i = 0; i < 3; ++i)
    {
        float col4[4] = {axes[3 * i], axes[3 * i + 1], axes[3 * i + 2], 0.f};
        float sq = (col4[0] * col4[0] + col4[1] * col4[1]) + (col4[2] * col4[2] + col4[3] * col4[3]);
        float scale = sqrtf(sq);
        obb[12 + i] = he[i] * scale;
    }
    return orc_intersect_obb(obb, o, d, dist);
}

/* ------------------------------------------------------------------------------------------------
 * ray - sphere: src/intersection.cpp:133-148 ; ray - plane: :10-21
 * ---------------------------------------------------------------------------------------------- */
int orc_intersect_sphere(const float *s4, const float *o, const float *d, float *dist)
{
    v3 l = sub(mk(s4), mk(o));
    v3 dir = mk(d);
    float s = dot(l, dir);
    float l2 = dot(l, l);
    float r2 = s4[3] * s4[3];
    if(s < 0 && l2 > r2) return 0;
    float m2 = l2 - s * s;
    if(m2 > r2) return 0;
    float q = sqrtf(r2 - m2);
    *dist = (l2 > r2) ? (s - q) : (s + q);
    return 1;
}

int orc_intersect_plane(const float *p4, const float *o, const float *d, float *dist)
{
    v3 nneg = {-p4[0], -p4[1], -p4[2]};
    float denom = dot(nneg, mk(d));
    if(fabs((double)denom) > 1e-6) /* std::abs(float) > 1e-6 (double literal): compared in double */
    {
        float dd = (p4[3] - dot(mk(o), nneg)) / denom;
        if(dd >= 0)
        {
            *dist = dd;
            return 1;
        }
    }
    return 0;
}

void orc_compute_aabb(const float *pts, uint32_t n, float *out6)
{
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for(uint32_t i = 0; i < n; ++i)
        for(int a = 0; a < 3; ++a)
        {
            float c = pts[3 * i + a];
            /* glm::min(x, y) = (y < x) ? y : x ; glm::max(x, y) = (x < y) ? y : x */
            mn[a] = (c < mn[a]) ? c : mn[a];
            mx[a] = (mx[a] < c) ? c : mx[a];
        }
    for(int a = 0; a < 3; ++a)
    {
        out6[a] = mn[a];
        out6[3 + a] = mx[a];
    }
}

/* compute_bounding_sphere (src/intersection.cpp:25-92, the `#if 1` branch): Ritter's sphere. Extremum
 * points per axis (strict comparisons: the first occurrence wins, and a NaN coordinate never replaces the
 * current candidate), the axis with the largest squared extent (strict >, from 0) seeds the sphere, then one
 * pass grows it to each point outside: k = 0.5 + (r / d) / 2, c = mix(p, c, k) = p (1 - k) + c k,
 * r = (r + d) / 2. glm::length2(v) = (x x + y y) + z z. out4 = center xyz, radius. n == 0 gives zeros
 * (the reference reads vertices[0] there). */
void orc_bounding_sphere(const float *pts, uint32_t n, float *out4)
{
    if(n == 0)
    {
        out4[0] = out4[1] = out4[2] = out4[3] = 0.f;
        return;
    }
    uint32_t pmin[3] = {0, 0, 0}, pmax[3] = {0, 0, 0};
    for(uint32_t i = 0; i < n; ++i)
        for(int a = 0; a < 3; ++a)
        {
            const float c = pts[3 * i + a];
            pmin[a] = (c < pts[3 * pmin[a] + a]) ? i : pmin[a];
            pmax[a] = (c > pts[3 * pmax[a] + a]) ? i : pmax[a];
        }
    float best = 0.f;
    int axis = 0;
    for(int a = 0; a < 3; ++a)
    {
        const float *p1 = pts + 3 * pmin[a], *p2 = pts + 3 * pmax[a];
        const float dx = p2[0] - p1[0], dy = p2[1] - p1[1], dz = p2[2] - p1[2];
        const float d2 = (dx * dx + dy * dy) + dz * dz;
        if(d2 > best)
        {
            best = d2;
            axis = a;
        }
    }
    const float *p1 = pts + 3 * pmin[axis], *p2 = pts + 3 * pmax[axis];
    float c[3] = {(p1[0] + p2[0]) / 2.f, (p1[1] + p2[1]) / 2.f, (p1[2] + p2[2]) / 2.f};
    float r = sqrtf(best) / 2.f;
    for(uint32_t i = 0; i < n; ++i)
    {
        const float *p = pts + 3 * i;
        const float vx = p[0] - c[0], vy = p[1] - c[1], vz = p[2] - c[2];
        const float d2 = (vx * vx + vy * vy) + vz * vz;
        if(d2 > r * r)
        {
            const float d = sqrtf(d2);
            const float k = 0.5f + (r / d) / 2;
            for(int a = 0; a < 3; ++a) c[a] = p[a] * (1.f - k) + c[a] * k;
            r = (r + d) / 2;
        }
    }
    out4[0] = c[0];
    out4[1] = c[1];
    out4[2] = c[2];
    out4[3] = r;
}

/* compute_centroid (src/intersection.cpp:110) returns crocore::mean(vertices); crocore is an empty
 * submodule in the reference checkout, so its summation order is unknown: this is the sequential mean
 * (sum from zero in index order, divided by n as float). Parity unpinned. */
void orc_centroid(const float *pts, uint32_t n, float *out3)
{
    float s[3] = {0.f, 0.f, 0.f};
    for(uint32_t i = 0; i < n; ++i)
        for(int a = 0; a < 3; ++a) s[a] = s[a] + pts[3 * i + a];
    for(int a = 0; a < 3; ++a) out3[a] = n ? s[a] / (float)n : 0.f;
}

/* ------------------------------------------------------------------------------------------------
 * brute-force closest / any hit (the harness the reference lacks; acceptance tmin < t < tmax,
 * ties on t resolve to the lowest triangle index)
 * ---------------------------------------------------------------------------------------------- */
typedef struct
{
    float t, u, v;
    uint32_t prim;
} hit16;

typedef struct
{
    const float *tris;
    uint32_t n_tri;
    const float *rays;
    uint32_t r0, r1;
    hit16 *hits;
    uint8_t *occ;
} job_t;

/* Culling criterion of the library's BVH traversal (not reference arithmetic: the library's conservative box
 * test, restated so that a test can decide whether a ray's brute-force winner could ever be culled).
 * Boxes: the triangle's vertex min/max grown by x -/+ |x| 2^-16 (bvh.hip pad_lo/pad_hi); every enclosing BVH
 * box contains this one. Slab: inv = 1/d (|d| < 1e-20 -> +-1e-20), plane t = fma(b, inv, c) with
 * c = -RN(o inv) -/+ 2^-20 |RN(o inv)| shifted toward smaller t on the entry plane, interval ends grown by
 * 2^-16 |t| with one fma each (traverse.hpp make_ray_setup / slab_interval). A box is entered iff
 * tn <= tf, tf >= tmin and tn <= best t. The slab values only move outward when the box grows, so when the
 * winner's own box is entered with best t = its t, no box holding the winner is ever culled and every BVH the
 * library builds returns the brute-force record. Returns 1 when the winner at t passes. */
static float cull_safe_dir(float x) { return fabsf(x) < 1e-20f ? copysignf(1e-20f, x) : x; }

int orc_cull_safe(const float *tri, const float *ray8, float t)
{
    float n_ = -INFINITY, f_ = INFINITY;
    for(int a = 0; a < 3; ++a)
    {
        const float x0 = tri[a], x1 = tri[3 + a], x2 = tri[6 + a];
        const float mn = fminf(fminf(x0, x1), x2), mx = fmaxf(fmaxf(x0, x1), x2);
        const float lo = mn - fabsf(mn) * 0x1p-16f, hi = mx + fabsf(mx) * 0x1p-16f;
        const float inv = 1.0f / cull_safe_dir(ray8[4 + a]);
        const float p = ray8[a] * inv;
        const float sh = fabsf(p) < INFINITY ? fabsf(p) * 0x1p-20f : 0.0f;
        const float clo = inv >= 0.0f ? -p - sh : -p + sh, chi = inv >= 0.0f ? -p + sh : -p - sh;
        const float s0 = fmaf(lo, inv, clo), s1 = fmaf(hi, inv, chi);
        n_ = fmaxf(n_, fminf(s0, s1));
        f_ = fminf(f_, fmaxf(s0, s1));
    }
    const float tn = fmaf(-fabsf(n_), 0x1p-16f, n_), tf = fmaf(fabsf(f_), 0x1p-16f, f_);
    return tn <= tf && tf >= ray8[3] && tn <= t;
}

static void closest_one(const float *tris, uint32_t n_tri, const float *r, hit16 *h)
{
    float best = INFINITY, bu = 0.f, bv = 0.f;
    uint32_t bp = 0xFFFFFFFFu;
    for(uint32_t i = 0; i < n_tri; ++i)
    {
        float t, u, v;
        if(orc_intersect_tri(tris + 9 * (size_t)i, r, r + 4, &t, &u, &v) && t > r[3] && t < r[7] &&
           (bp == 0xFFFFFFFFu || t < best))
        {
            best = t;
            bu = u;
            bv = v;
            bp = i;
        }
    }
    h->t = (bp == 0xFFFFFFFFu) ? INFINITY : best;
    h->u = bu;
    h->v = bv;
    h->prim = bp;
}

static void *closest_job(void *arg)
{
    job_t *j = (job_t *)arg;
    for(uint32_t r = j->r0; r < j->r1; ++r)
    {
        const float *ray = j->rays + 8 * (size_t)r;
        closest_one(j->tris, j->n_tri, ray, j->hits + r);
        if(j->occ)// the culling criterion of the winner (1 for a miss)
            j->occ[r] = j->hits[r].prim == 0xFFFFFFFFu
                                ? 1
                                : (uint8_t)orc_cull_safe(j->tris + 9 * (size_t)j->hits[r].prim, ray, j->hits[r].t);
    }
    return NULL;
}

static void *any_job(void *arg)
{
    job_t *j = (job_t *)arg;
    for(uint32_t r = j->r0; r < j->r1; ++r)
    {
        const float *ray = j->rays + 8 * (size_t)r;
        uint8_t occ = 0;
        for(uint32_t i = 0; i < j->n_tri && !occ; ++i)
        {
            float t, u, v;
            if(orc_intersect_tri(j->tris + 9 * (size_t)i, ray, ray + 4, &t, &u, &v) && t > ray[3] && t < ray[7])
                occ = 1;
        }
        j->occ[r] = occ;
    }
    return NULL;
}

static void run_jobs(void *(*fn)(void *), const float *tris, uint32_t n_tri, const float *rays, uint32_t n_ray,
                     hit16 *hits, uint8_t *occ, int nthreads)
{
    if(nthreads < 1) nthreads = 1;
    if((uint32_t)nthreads > n_ray) nthreads = n_ray ? (int)n_ray : 1;
    pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
    job_t *jobs = (job_t *)calloc((size_t)nthreads, sizeof(job_t));
    for(int k = 0; k < nthreads; ++k)
    {
        jobs[k] = (job_t){tris, n_tri, rays, (uint32_t)((uint64_t)n_ray * k / nthreads),
                          (uint32_t)((uint64_t)n_ray * (k + 1) / nthreads), hits, occ};
        if(nthreads == 1) fn(&jobs[k]);
        else pthread_create(&th[k], NULL, fn, &jobs[k]);
    }
    if(nthreads > 1)
        for(int k = 0; k < nthreads; ++k) pthread_join(th[k], NULL);
    free(th);
    free(jobs);
}

void orc_brute_closest(const float *tris, uint32_t n_tri, const float *rays8, uint32_t n_ray, void *hits16,
                       int nthreads)
{
    run_jobs(closest_job, tris, n_tri, rays8, n_ray, (hit16 *)hits16, NULL, nthreads);
}

void orc_brute_closest_cull_safe(const float *tris, uint32_t n_tri, const float *rays8, uint32_t n_ray,
                                 void *hits16, uint8_t *safe, int nthreads)
{
    run_jobs(closest_job, tris, n_tri, rays8, n_ray, (hit16 *)hits16, safe, nthreads);
}

void orc_brute_any(const float *tris, uint32_t n_tri, const float *rays8, uint32_t n_ray, uint8_t *occluded,
                   int nthreads)
{
    run_jobs(any_job, tris, n_tri, rays8, n_ray, NULL, occluded, nthreads);
}

/* Two-level scene (the instance level of RayBuilder's top-level structure, src/RayBuilder.cpp:386-781):
 * each instance has a row-major 3x4 world-to-object matrix W (the reference's entry_t::inv_transform) and a
 * mesh. The object-space ray is o' = W (o, 1), d' = W (d, 0), each row as ((w0 x + w1 y) + w2 z) + w3,
 * with d' not normalized, so t is the same parameter in both spaces. A candidate counts iff tmin < t < tmax;
 * the smallest t wins, then the lowest instance, then the lowest triangle of that instance's mesh.
 * tris: all meshes' flat triangles (9 floats each) concatenated; tri_offset[m] .. tri_offset[m + 1] is mesh
 * m. out: 8 words per ray (t, u, v, prim, instance, custom index, 0, 0). */
void orc_scene_closest(const float *tris, const uint32_t *tri_offset, const float *w2o, const uint32_t *inst_mesh,
                       const uint32_t *inst_custom, uint32_t n_inst, const float *rays8, uint32_t n_ray, uint32_t *out8)
{
    for(uint32_t r = 0; r < n_ray; ++r)
    {
        const float *ray = rays8 + 8 * (size_t)r;
        float best = INFINITY, bu = 0.f, bv = 0.f;
        uint32_t bp = 0xFFFFFFFFu, bi = 0xFFFFFFFFu;
        for(uint32_t k = 0; k < n_inst; ++k)
        {
            const float *w = w2o + 12 * (size_t)k;
            float o[3], d[3];
            for(int a = 0; a < 3; ++a)
            {
                const float *row = w + 4 * a;
                o[a] = ((row[0] * ray[0] + row[1] * ray[1]) + row[2] * ray[2]) + row[3];
                d[a] = (row[0] * ray[4] + row[1] * ray[5]) + row[2] * ray[6];
            }
            const uint32_t m = inst_mesh[k];
            for(uint32_t i = tri_offset[m]; i < tri_offset[m + 1]; ++i)
            {
                float t, u, v;
                if(orc_intersect_tri(tris + 9 * (size_t)i, o, d, &t, &u, &v) && t > ray[3] && t < ray[7] &&
                   (bp == 0xFFFFFFFFu || t < best))
                {
                    best = t;
                    bu = u;
                    bv = v;
                    bp = i - tri_offset[m];
                    bi = k;
                }
            }
        }
        uint32_t *h = out8 + 8 * (size_t)r;
        const float tt = bp == 0xFFFFFFFFu ? INFINITY : best;
        memcpy(h, &tt, 4);
        memcpy(h + 1, &bu, 4);
        memcpy(h + 2, &bv, 4);
        h[3] = bp;
        h[4] = bi;
        h[5] = bi == 0xFFFFFFFFu ? 0u : inst_custom[bi];
        h[6] = h[7] = 0u;
    }
}

/* transmittance_test (shaders/slang/ray/visibility_test.slang:94-165) by brute force: 0 if a triangle with
 * transmission == 0 is hit in (tmin, tmax); otherwise the product of transmission * color (per channel) over
 * the transmissive hits, multiplied in ascending (t, triangle) order. With `cull`, triangles whose
 * det = dot(e1, cross(d, e2)) is negative are skipped unless two-sided. Per triangle: trans[i], color[3i..],
 * two_sided[i]. out4: rgb, 0 per ray. */
typedef struct
{
    float t;
    uint32_t prim;
} cand_t;

static int cand_cmp(const void *a, const void *b)
{
    const cand_t *x = (const cand_t *)a, *y = (const cand_t *)b;
    if(x->t < y->t) return -1;
    if(x->t > y->t) return 1;
    return x->prim < y->prim ? -1 : (x->prim > y->prim ? 1 : 0);
}

void orc_transmittance(const float *tris, const float *trans, const float *color, const uint8_t *two_sided,
                       uint32_t n_tri, const float *rays8, uint32_t n_ray, int cull, float *out4)
{
    cand_t *c = (cand_t *)malloc(sizeof(cand_t) * (n_tri ? n_tri : 1));
    for(uint32_t r = 0; r < n_ray; ++r)
    {
        const float *ray = rays8 + 8 * (size_t)r;
        uint32_t nc = 0;
        int blocked = 0;
        for(uint32_t i = 0; i < n_tri && !blocked; ++i)
        {
            const float *tr = tris + 9 * (size_t)i;
            float t, u, v;
            if(!orc_intersect_tri(tr, ray, ray + 4, &t, &u, &v) || !(t > ray[3] && t < ray[7])) continue;
            if(cull && !two_sided[i])
            {
                const float e1[3] = {tr[3] - tr[0], tr[4] - tr[1], tr[5] - tr[2]};
                const float e2[3] = {tr[6] - tr[0], tr[7] - tr[1], tr[8] - tr[2]};
                const float *d = ray + 4;
                const float p[3] = {d[1] * e2[2] - e2[1] * d[2], d[2] * e2[0] - e2[2] * d[0], d[0] * e2[1] - e2[0] * d[1]};
                const float det = (e1[0] * p[0] + e1[1] * p[1]) + e1[2] * p[2];
                if(det < 0.f) continue;
            }
            if(!(trans[i] > 0.f))
            {
                blocked = 1;
                break;
            }
            c[nc].t = t;
            c[nc].prim = i;
            ++nc;
        }
        float T[3] = {1.f, 1.f, 1.f};
        if(blocked) T[0] = T[1] = T[2] = 0.f;
        else
        {
            qsort(c, nc, sizeof(cand_t), cand_cmp);
            for(uint32_t k = 0; k < nc; ++k)
                for(int a = 0; a < 3; ++a) T[a] = T[a] * (trans[c[k].prim] * color[3 * (size_t)c[k].prim + a]);
        }
        float *o = out4 + 4 * (size_t)r;
        o[0] = T[0];
        o[1] = T[1];
        o[2] = T[2];
        o[3] = 0.f;
    }
    free(c);
}

uint32_t orc_closest_f64(const float *tris, uint32_t n_tri, const float *r, double *bt, double *bu, double *bv)
{
    double best = INFINITY;
    uint32_t bp = 0xFFFFFFFFu;
    *bu = *bv = 0.0;
    for(uint32_t i = 0; i < n_tri; ++i)
    {
        double t, u, v;
        if(orc_intersect_tri_f64(tris + 9 * (size_t)i, r, r + 4, &t, &u, &v) && t > r[3] && t < r[7] &&
           (bp == 0xFFFFFFFFu || t < best))
        {
            best = t;
            *bu = u;
            *bv = v;
            bp = i;
        }
    }
    *bt = best;
    return bp;
}

/* ------------------------------------------------------------------------------------------------
 * Geometry::IcoSphere (src/Geometry.cpp:566-613) with tessellate (:141-192)
 * ---------------------------------------------------------------------------------------------- */
void orc_icosphere_size(uint32_t subdiv, uint32_t *n_vert, uint32_t *n_idx)
{
    uint32_t nv = 12, nt = 20;
    for(uint32_t i = 0; i < subdiv; ++i)
    {
        nv += 3 * nt; /* three new vertex copies per triangle (:157-171) */
        nt *= 4;
    }
    *n_vert = nv;
    *n_idx = 3 * nt;
}

void orc_icosphere(float radius, uint32_t subdiv, float *pos, uint32_t *idx)
{
    const float phi = 1.61803398874989484820f; /* std::numbers::phi_v<float> */
    const float e1 = radius / sqrtf(1.f + phi * phi);
    const float e2 = phi * e1;
    const float base[12][3] = {{-e1, e2, 0},  {e1, e2, 0},  {-e1, -e2, 0}, {e1, -e2, 0},
                               {0, -e1, e2},  {0, e1, e2},  {0, -e1, -e2}, {0, e1, -e2},
                               {e2, 0, -e1},  {e2, 0, e1},  {-e2, 0, -e1}, {-e2, 0, e1}};
    const uint32_t base_idx[60] = {0, 11, 5, 0, 5,  1,  0,  1,  7,  0,  7,  10, 0, 10, 11, 1, 5, 9, 5, 11,
                                   4, 11, 10, 2, 10, 7,  6,  7,  1,  8,  3,  9,  4, 3,  4,  2, 3, 2, 6, 3,
                                   6, 8,  3,  8, 9,  4,  9,  5,  2,  4,  11, 6,  2, 10, 8,  6, 7, 9, 8, 1};
    uint32_t nv = 12, ni = 60;
    memcpy(pos, base, sizeof(base));
    uint32_t tmp_n, total_idx;
    orc_icosphere_size(subdiv, &tmp_n, &total_idx);
    uint32_t *cur = (uint32_t *)malloc(sizeof(uint32_t) * total_idx);
    uint32_t *nxt = (uint32_t *)malloc(sizeof(uint32_t) * total_idx);
    memcpy(cur, base_idx, sizeof(base_idx));
    for(uint32_t level = 0; level < subdiv; ++level)
    {
        uint32_t nn = 0;
        for(uint32_t j = 0; j < ni; j += 3)
        {
            uint32_t k = nv;
            uint32_t a = cur[j], b = cur[j + 1], c = cur[j + 2];
            const uint32_t quad[12] = {a, k + 1, k, k + 1, b, k + 2, k, k + 2, c, k, k + 1, k + 2}; /* :160-161 */
            memcpy(nxt + nn, quad, sizeof(quad));
            nn += 12;
            /* the lerped midpoints (:163-171) are overwritten by the tessellation control of IcoSphere
             * (:588-596): ab = k+1, ac = k, bc = k+2, each radius * normalize(va + vb) */
            v3 va = mk(pos + 3 * a), vb = mk(pos + 3 * b), vc = mk(pos + 3 * c);
            v3 ab = normalize(add(va, vb)), ac = normalize(add(va, vc)), bc = normalize(add(vb, vc));
            ab = (v3){radius * ab.x, radius * ab.y, radius * ab.z};
            ac = (v3){radius * ac.x, radius * ac.y, radius * ac.z};
            bc = (v3){radius * bc.x, radius * bc.y, radius * bc.z};
            float *p = pos + 3 * (size_t)k;
            p[0] = ac.x; p[1] = ac.y; p[2] = ac.z;
            p[3] = ab.x; p[4] = ab.y; p[5] = ab.z;
            p[6] = bc.x; p[7] = bc.y; p[8] = bc.z;
            nv += 3;
        }
        ni = nn;
        uint32_t *sw = cur;
        cur = nxt;
        nxt = sw;
    }
    memcpy(idx, cur, sizeof(uint32_t) * ni);
    free(cur);
    free(nxt);
}

/* Geometry::Plane(width, height, numSegments_W, numSegments_H, xy_plane=false), src/Geometry.cpp:313-371 */
void orc_plane(float width, float height, uint32_t sw, uint32_t sh, float *pos, uint32_t *idx)
{
    float width_half = width / 2, height_half = height / 2;
    float segment_width = width / sw, segment_height = height / sh;
    uint32_t gx1 = sw + 1, gz1 = sh + 1;
    size_t v = 0;
    for(uint32_t iz = 0; iz < gz1; ++iz)
        for(uint32_t ix = 0; ix < gx1; ++ix)
        {
            float x = ix * segment_width - width_half;
            float y = iz * segment_height - height_half;
            pos[3 * v + 0] = x;
            pos[3 * v + 1] = 0;
            pos[3 * v + 2] = y;
            ++v;
        }
    size_t n = 0;
    for(uint32_t iz = 0; iz < sh; ++iz)
        for(uint32_t ix = 0; ix < sw; ++ix)
        {
            uint32_t a = ix + gx1 * iz, b = ix + gx1 * (iz + 1), c = (ix + 1) + gx1 * (iz + 1), d = (ix + 1) + gx1 * iz;
            idx[n++] = a; idx[n++] = b; idx[n++] = c;
            idx[n++] = c; idx[n++] = d; idx[n++] = a;
        }
}

/* ------------------------------------------------------------------------------------------------
 * RNG of the path tracer
 * ---------------------------------------------------------------------------------------------- */
uint32_t orc_xxhash32(uint32_t lhs, uint32_t rhs) /* shaders/slang/utils/hash.slang:9-18 */
{
    const uint32_t P2 = 2246822519U, P3 = 3266489917U, P4 = 668265263U, P5 = 374761393U;
    uint32_t h = lhs + P5 + rhs * P3;
    h = P4 * ((h << 17) | (h >> 15));
    h = P2 * (h ^ (h >> 15));
    h = P3 * (h ^ (h >> 13));
    return h ^ (h >> 16);
}

uint32_t orc_lcg(uint32_t *s) /* shaders/slang/utils/random.slang:27-33 */
{
    *s = 1664525u * *s + 1013904223u;
    return *s & 0x00FFFFFFu;
}

float orc_rnd(uint32_t *s) { return (float)orc_lcg(s) / (float)0x01000000; } /* random.slang:36 */

static uint32_t bitrev32(uint32_t x)
{
    uint32_t r = 0;
    for(int i = 0; i < 32; ++i) r |= ((x >> i) & 1u) << (31 - i);
    return r;
}

void orc_hammersley(uint32_t i, uint32_t n, float *out2) /* shaders/slang/utils/sampling.slang:9-13 */
{
    /* an unsuffixed Slang float literal is a float: 2.3283064365386963e-10f == 2^-32 exactly */
    out2[0] = (float)i / (float)n;
    out2[1] = (float)bitrev32(i) * 2.3283064365386963e-10f;
}
