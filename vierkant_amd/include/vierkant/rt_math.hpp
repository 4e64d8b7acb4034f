// rt_math.hpp - reference arithmetic of vierkant's intersection routines, shared by the host
// API (vierkant/intersection.hpp) and the gfx950 kernels.
//
// Every function reproduces the operation order of the reference (glm semantics):
//   dot(a, b)    = (a.x*b.x + a.y*b.y) + a.z*b.z
//   cross(x, y)  = (x.y*y.z - y.y*x.z, x.z*y.x - y.z*x.x, x.x*y.y - y.x*x.y)
//   normalize(v) = v * (1 / sqrt(dot(v, v)))
// and must be compiled with -ffp-contract=off (no fused multiply-add) and IEEE division and
// square root, so that host and device produce the same bits as the reference's x86-64 build.
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#define VK_HD __host__ __device__ __forceinline__
#else
#define VK_HD inline
#endif

namespace vierkant::rt
{

struct vec3
{
    float x, y, z;
};

VK_HD vec3 make_vec3(float x, float y, float z) { return vec3{x, y, z}; }
VK_HD vec3 operator-(const vec3 &a, const vec3 &b) { return vec3{a.x - b.x, a.y - b.y, a.z - b.z}; }
VK_HD vec3 operator+(const vec3 &a, const vec3 &b) { return vec3{a.x + b.x, a.y + b.y, a.z + b.z}; }
VK_HD vec3 operator*(const vec3 &a, float s) { return vec3{a.x * s, a.y * s, a.z * s}; }
VK_HD vec3 operator*(float s, const vec3 &a) { return vec3{s * a.x, s * a.y, s * a.z}; }
VK_HD vec3 operator-(const vec3 &a) { return vec3{-a.x, -a.y, -a.z}; }

VK_HD float dot(const vec3 &a, const vec3 &b)
{
    float mx = a.x * b.x, my = a.y * b.y, mz = a.z * b.z;
    return (mx + my) + mz;
}

VK_HD vec3 cross(const vec3 &x, const vec3 &y)
{
    return vec3{x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y};
}

VK_HD float fabs_(float x) { return __builtin_fabsf(x); }

constexpr uint32_t REJECT = 0, INTERSECT = 1, INSIDE = 2;

//! Möller–Trumbore exactly as src/intersection.cpp:114-129, on a triangle given as (v0, e1, e2)
//! with e1 = v1 - v0 and e2 = v2 - v0 (the reference's first line, :116, so precomputing them
//! changes no bit). Returns true on INTERSECT; t/u/v written only then.
VK_HD bool mt_reference(const vec3 &o, const vec3 &d, const vec3 &v0, const vec3 &e1, const vec3 &e2, float &t,
                        float &u, float &v)
{
    vec3 pvec = cross(d, e2);
    float det = dot(e1, pvec);
    constexpr float epsilon = 10.0e-10f;
    if(det > -epsilon && det < epsilon) return false;
    float inv_det = 1.0f / det;
    vec3 tvec = o - v0;
    float uu = inv_det * dot(tvec, pvec);
    if(uu < 0.0f || uu > 1.0f) return false;
    vec3 qvec = cross(tvec, e1);
    float vv = dot(d, qvec) * inv_det;
    if(vv < 0.0f || (uu + vv) > 1.0f) return false;
    t = dot(e2, qvec) * inv_det;
    u = uu;
    v = vv;
    return true;
}

//! Same decision and the same t/u/v bits as mt_reference, but the division 1/det is only
//! evaluated when division-free bounds on a = dot(tvec, pvec) and b = dot(d, qvec) cannot decide.
//!
//! With D = |det|, A = a*sign(det), B = b*sign(det) (exact), and D in [1e-9, 2^100]:
//!   A <= -D*2^-80            => u = RN(RN(1/det)*a) is a negative normal float      => REJECT
//!   A >= RN(D*(1+2^-19))     => u >= (1+2^-20)(1-2^-24)^2 > 1                        => REJECT
//!   B <= -D*2^-80            => v < 0                                                => REJECT
//!   RN(A+B) >= RN(D*(1+2^-16)) (given A, B > -D*2^-80) => RN(u+v) >= 1+2^-18 > 1     => REJECT
//! (u and v each carry at most two roundings, relative error < 2^-23.) NaN or infinite
//! intermediates fail every bound and fall through to the reference sequence.
//! With `cull_back` set, a triangle whose det is negative (its counter-clockwise normal faces away from
//! the ray) is rejected as well; that is the back-face culling of a ray traced in air.
//! With WINDOW set, a hit is reported only for t_above < t < t_below (a closest-hit loop's window), a test that
//! then joins the tail's last branch.
template<bool WINDOW>
VK_HD bool mt_fast_impl(const vec3 &o, const vec3 &d, const vec3 &v0, const vec3 &e1, const vec3 &e2, float &t,
                        float &u, float &v, bool cull_back, float t_above, float t_below)
{
    // every division-free test is evaluated and joined with '|' into one branch: on a GPU the early-outs
    // cost more in divergent branching than the few products they skip
    vec3 pvec = cross(d, e2);
    float det = dot(e1, pvec);
    constexpr float epsilon = 10.0e-10f;
    vec3 tvec = o - v0;
    float a = dot(tvec, pvec);
    vec3 qvec = cross(tvec, e1);
    float b = dot(d, qvec);
    const bool neg = det < 0.f;
    const float D = neg ? -det : det;
    const bool bounded = D <= 0x1p100f;
    const float lo = D * 0x1p-80f;
    const float A = neg ? -a : a;
    const float B = neg ? -b : b;
    const bool reject = ((det > -epsilon) & (det < epsilon)) | (cull_back & neg) |
                        (bounded & ((A <= -lo) | (A >= D * (1.0f + 0x1p-19f)) | (B <= -lo) |
                                    ((A + B) >= D * (1.0f + 0x1p-16f))));
    if(reject) return false;
    // reference tail (src/intersection.cpp:121-128)
    float inv_det = 1.0f / det;
    float uu = inv_det * a;
    if(uu < 0.0f || uu > 1.0f) return false;
    float vv = b * inv_det;
    if(vv < 0.0f || (uu + vv) > 1.0f) return false;
    const float tt = dot(e2, qvec) * inv_det;
    if(WINDOW && !(tt > t_above && tt < t_below)) return false;
    t = tt;
    u = uu;
    v = vv;
    return true;
}

VK_HD bool mt_fast(const vec3 &o, const vec3 &d, const vec3 &v0, const vec3 &e1, const vec3 &e2, float &t, float &u,
                   float &v, bool cull_back = false)
{
    return mt_fast_impl<false>(o, d, v0, e1, e2, t, u, v, cull_back, 0.0f, 0.0f);
}

//! mt_fast's hit restricted to t_above < t < t_below
VK_HD bool mt_fast_window(const vec3 &o, const vec3 &d, const vec3 &v0, const vec3 &e1, const vec3 &e2, float t_above,
                          float t_below, float &t, float &u, float &v)
{
    return mt_fast_impl<true>(o, d, v0, e1, e2, t, u, v, false, t_above, t_below);
}

//! intersect(const OBB&, const Ray&), src/intersection.cpp:152-184. axis[i] are the OBB's columns.
VK_HD uint32_t obb_reference(const vec3 &center, const vec3 axis[3], const vec3 &half_lengths, const vec3 &o,
                             const vec3 &d, float &distance)
{
    constexpr float epsilon = 1e-10f;
    float t_min = -3.40282347e+38f;
    float t_max = 3.40282347e+38f;
    vec3 p = center - o;
    const float h[3] = {half_lengths.x, half_lengths.y, half_lengths.z};
    for(int i = 0; i < 3; i++)
    {
        float e = dot(axis[i], p);
        float f = dot(axis[i], d);
        if(fabs_(f) > epsilon)
        {
            float t1 = (e + h[i]) / f;
            float t2 = (e - h[i]) / f;
            if(t1 > t2)
            {
                float tmp = t1;
                t1 = t2;
                t2 = tmp;
            }
            if(t1 > t_min) t_min = t1;
            if(t2 < t_max) t_max = t2;
            if(t_min > t_max) return REJECT;
            if(t_max < 0) return REJECT;
        }
        else if((-e - h[i]) > 0 || (-e + h[i]) < 0) { return REJECT; }
    }
    distance = (t_min > 0) ? t_min : t_max;
    return INTERSECT;
}

//! AABB::intersect(const Ray&), src/intersection.cpp:308-312: OBB(aabb, mat4(1)) (:219-227) then the
//! slab test. With the identity transform each center component is (m0*x + m1*y) + (m2*z + m3*1)
//! with 0/1 entries (glm's mat4*vec4), axes are unit vectors and scale is 1.
VK_HD uint32_t aabb_reference(const vec3 &mn, const vec3 &mx, const vec3 &o, const vec3 &d, float &distance)
{
    vec3 c{(mx.x + mn.x) / 2.f, (mx.y + mn.y) / 2.f, (mx.z + mn.z) / 2.f};
    vec3 he{(mx.x - mn.x) / 2.f, (mx.y - mn.y) / 2.f, (mx.z - mn.z) / 2.f};
    vec3 center{(1.f * c.x + 0.f * c.y) + (0.f * c.z + 0.f), (0.f * c.x + 1.f * c.y) + (0.f * c.z + 0.f),
                (0.f * c.x + 0.f * c.y) + (1.f * c.z + 0.f)};
    const vec3 axis[3] = {{1.f, 0.f, 0.f}, {0.f, 1.f, 0.f}, {0.f, 0.f, 1.f}};
    return obb_reference(center, axis, he, o, d, distance);
}

//! intersect(const Sphere&, const Ray&), src/intersection.cpp:133-148
VK_HD uint32_t sphere_reference(const vec3 &center, float radius, const vec3 &o, const vec3 &d, float &distance)
{
    vec3 l = center - o;
    float s = dot(l, d);
    float l2 = dot(l, l);
    float r2 = radius * radius;
    if(s < 0 && l2 > r2) return REJECT;
    float m2 = l2 - s * s;
    if(m2 > r2) return REJECT;
    float q = sqrtf(r2 - m2);
    distance = (l2 > r2) ? (s - q) : (s + q);
    return INTERSECT;
}

//! intersect(const Plane&, const Ray&), src/intersection.cpp:10-21 (coefficients a, b, c, d)
VK_HD uint32_t plane_reference(const float coeff[4], const vec3 &o, const vec3 &d, float &distance)
{
    vec3 nneg{-coeff[0], -coeff[1], -coeff[2]};
    float denom = dot(nneg, d);
    // std::abs(float) > 1e-6: the float is promoted and compared against the double literal
    if((double)fabs_(denom) > 1e-6)
    {
        float dd = (coeff[3] - dot(o, nneg)) / denom;
        if(dd >= 0)
        {
            distance = dd;
            return INTERSECT;
        }
    }
    return REJECT;
}

// ------------------------------------------------------------------------------------------------
// triangle / AABB overlap: separating-axis test (Akenine-Moller 2001) as intersect(const Triangle&,
// const AABB&) evaluates it (src/intersection.cpp:364-370 -> src/triangle_intersection.h:753-843):
// 9 edge x axis tests, the 3 box axes, then the triangle plane. Each edge test projects the two vertices
// the reference projects (the third projects like one of them in exact arithmetic).
// ------------------------------------------------------------------------------------------------
namespace detail
{
struct v3a
{
    float c[3];
};

//! one edge x axis separating test: projections p_a, p_b of two vertices on axis (a, b) in the plane
//! (i, j), radius fa h[i] + fb h[j]; `second_first` swaps the comparison order used for min/max
VK_HD bool axis_separates(float a, float b, float fa, float fb, const float *va, const float *vb, int i, int j,
                          bool neg_first, bool second_first, const float h[3])
{
    float pa, pb;
    if(!neg_first)
    {
        pa = a * va[i] - b * va[j];
        pb = a * vb[i] - b * vb[j];
    }
    else
    {
        pa = -a * va[i] + b * va[j];
        pb = -a * vb[i] + b * vb[j];
    }
    float mn, mx;
    if(!second_first)
    {
        if(pa < pb) mn = pa, mx = pb;
        else mn = pb, mx = pa;
    }
    else
    {
        if(pb < pa) mn = pb, mx = pa;
        else mn = pa, mx = pb;
    }
    float rad = fa * h[i] + fb * h[j];
    return mn > rad || mx < -rad;
}

VK_HD void minmax3(float x0, float x1, float x2, float &mn, float &mx)
{
    mn = mx = x0;
    if(x1 < mn) mn = x1;
    if(x1 > mx) mx = x1;
    if(x2 < mn) mn = x2;
    if(x2 > mx) mx = x2;
}
}// namespace detail

VK_HD uint32_t tri_box_overlap(const float tri[9], const float box_min[3], const float box_max[3])
{
    float center[3], h[3], v0[3], v1[3], v2[3], e0[3], e1[3], e2[3];
    for(int k = 0; k < 3; ++k)
    {
        center[k] = (box_max[k] + box_min[k]) / 2.f;// AABB::center()
        h[k] = (box_max[k] - box_min[k]) / 2.f;     // AABB::half_extents()
    }
    for(int k = 0; k < 3; ++k)
    {
        v0[k] = tri[k] - center[k];
        v1[k] = tri[3 + k] - center[k];
        v2[k] = tri[6 + k] - center[k];
    }
    for(int k = 0; k < 3; ++k)
    {
        e0[k] = v1[k] - v0[k];
        e1[k] = v2[k] - v1[k];
        e2[k] = v0[k] - v2[k];
    }
    using detail::axis_separates;
    // edge e0: x-axis (v0, v2), y-axis (v0, v2), z-axis (v1, v2 compared second-first)
    float fx = fabs_(e0[0]), fy = fabs_(e0[1]), fz = fabs_(e0[2]);
    if(axis_separates(e0[2], e0[1], fz, fy, v0, v2, 1, 2, false, false, h)) return 0;
    if(axis_separates(e0[2], e0[0], fz, fx, v0, v2, 0, 2, true, false, h)) return 0;
    if(axis_separates(e0[1], e0[0], fy, fx, v1, v2, 0, 1, false, true, h)) return 0;
    // edge e1: x (v0, v2), y (v0, v2), z (v0, v1)
    fx = fabs_(e1[0]), fy = fabs_(e1[1]), fz = fabs_(e1[2]);
    if(axis_separates(e1[2], e1[1], fz, fy, v0, v2, 1, 2, false, false, h)) return 0;
    if(axis_separates(e1[2], e1[0], fz, fx, v0, v2, 0, 2, true, false, h)) return 0;
    if(axis_separates(e1[1], e1[0], fy, fx, v0, v1, 0, 1, false, false, h)) return 0;
    // edge e2: x (v0, v1), y (v0, v1), z (v1, v2 compared second-first)
    fx = fabs_(e2[0]), fy = fabs_(e2[1]), fz = fabs_(e2[2]);
    if(axis_separates(e2[2], e2[1], fz, fy, v0, v1, 1, 2, false, false, h)) return 0;
    if(axis_separates(e2[2], e2[0], fz, fx, v0, v1, 0, 2, true, false, h)) return 0;
    if(axis_separates(e2[1], e2[0], fy, fx, v1, v2, 0, 1, false, true, h)) return 0;
    // the three box axes
    for(int k = 0; k < 3; ++k)
    {
        float mn, mx;
        detail::minmax3(v0[k], v1[k], v2[k], mn, mx);
        if(mn > h[k] || mx < -h[k]) return 0;
    }
    // the triangle's plane against the box (planeBoxOverlap)
    float n[3] = {e0[1] * e1[2] - e0[2] * e1[1], e0[2] * e1[0] - e0[0] * e1[2], e0[0] * e1[1] - e0[1] * e1[0]};
    float vmin[3], vmax[3];
    for(int q = 0; q < 3; ++q)
    {
        float v = v0[q];
        if(n[q] > 0.0f)
        {
            vmin[q] = -h[q] - v;
            vmax[q] = h[q] - v;
        }
        else
        {
            vmin[q] = h[q] - v;
            vmax[q] = -h[q] - v;
        }
    }
    if((n[0] * vmin[0] + n[1] * vmin[1]) + n[2] * vmin[2] > 0.0f) return 0;
    if((n[0] * vmax[0] + n[1] * vmax[1]) + n[2] * vmax[2] >= 0.0f) return 1;
    return 0;
}

// ------------------------------------------------------------------------------------------------
// triangle / triangle overlap: orientation-predicate test of Guigue and Devillers (JGT 2003) as
// intersect(const Triangle&, const Triangle&) evaluates it (src/intersection.cpp:374-378 ->
// src/triangle_intersection.h:145-297, 2D :580-636): plane-side signs with a 1e-12 zero band, the
// canonical vertex permutations, two interval checks with ">= -eps", and the 2D test for coplanar pairs
// on the projection of largest area. Boundaries count as overlap.
// ------------------------------------------------------------------------------------------------
namespace gd
{
struct p2
{
    float x, y;
};
constexpr float EPS = 1e-12f;

VK_HD float orient(const p2 &a, const p2 &b, const p2 &c) { return (a.x - c.x) * (b.y - c.y) - (a.y - c.y) * (b.x - c.x); }

//! the "vertex" configuration of the 2D test (P1 is a vertex of T1 outside T2's vertex region)
VK_HD bool vertex_case(p2 P1, p2 Q1, p2 R1, p2 P2, p2 Q2, p2 R2)
{
    if(orient(R2, P2, Q1) >= 0.f)
    {
        if(orient(R2, Q2, Q1) <= 0.f)
        {
            if(orient(P1, P2, Q1) > 0.f) return orient(P1, Q2, Q1) <= 0.f;
            if(orient(P1, P2, R1) >= 0.f) return orient(Q1, R1, P2) >= 0.f;
            return false;
        }
        if(orient(P1, Q2, Q1) <= 0.f && orient(R2, Q2, R1) <= 0.f) return orient(Q1, R1, Q2) >= 0.f;
        return false;
    }
    if(orient(R2, P2, R1) >= 0.f)
    {
        if(orient(Q1, R1, R2) >= 0.f) return orient(P1, P2, R1) >= 0.f;
        if(orient(Q1, R1, Q2) >= 0.f) return orient(R2, R1, Q2) >= 0.f;
    }
    return false;
}

//! the "edge" configuration of the 2D test
VK_HD bool edge_case(p2 P1, p2 Q1, p2 R1, p2 P2, p2 Q2, p2 R2)
{
    (void)Q2;
    if(orient(R2, P2, Q1) >= 0.f)
    {
        if(orient(P1, P2, Q1) >= 0.f) return orient(P1, Q1, R2) >= 0.f;
        if(orient(Q1, R1, P2) >= 0.f) return orient(R1, P1, P2) >= 0.f;
        return false;
    }
    if(orient(R2, P2, R1) >= 0.f && orient(P1, P2, R1) >= 0.f)
    {
        if(orient(P1, R1, R2) >= 0.f) return true;
        return orient(Q1, R1, R2) >= 0.f;
    }
    return false;
}

//! both triangles counter-clockwise
VK_HD bool ccw_2d(p2 p1, p2 q1, p2 r1, p2 p2_, p2 q2, p2 r2)
{
    if(orient(p2_, q2, p1) >= 0.f)
    {
        if(orient(q2, r2, p1) >= 0.f)
        {
            if(orient(r2, p2_, p1) >= 0.f) return true;
            return edge_case(p1, q1, r1, p2_, q2, r2);
        }
        if(orient(r2, p2_, p1) >= 0.f) return edge_case(p1, q1, r1, r2, p2_, q2);
        return vertex_case(p1, q1, r1, p2_, q2, r2);
    }
    if(orient(q2, r2, p1) >= 0.f)
    {
        if(orient(r2, p2_, p1) >= 0.f) return edge_case(p1, q1, r1, q2, r2, p2_);
        return vertex_case(p1, q1, r1, q2, r2, p2_);
    }
    return vertex_case(p1, q1, r1, r2, p2_, q2);
}

VK_HD bool overlap_2d(p2 p1, p2 q1, p2 r1, p2 p2_, p2 q2, p2 r2)
{
    const bool cw1 = orient(p1, q1, r1) < 0.f, cw2 = orient(p2_, q2, r2) < 0.f;
    if(cw1) return cw2 ? ccw_2d(p1, r1, q1, p2_, r2, q2) : ccw_2d(p1, r1, q1, p2_, q2, r2);
    return cw2 ? ccw_2d(p1, q1, r1, p2_, r2, q2) : ccw_2d(p1, q1, r1, p2_, q2, r2);
}

VK_HD vec3 ld(const float *p) { return vec3{p[0], p[1], p[2]}; }

//! coplanar triangles: project on the axis-aligned plane where N1 has its largest component
VK_HD bool coplanar(const vec3 &p1, const vec3 &q1, const vec3 &r1, const vec3 &p2_, const vec3 &q2, const vec3 &r2,
                    const vec3 &n1)
{
    const float nx = n1.x < 0.f ? -n1.x : n1.x, ny = n1.y < 0.f ? -n1.y : n1.y, nz = n1.z < 0.f ? -n1.z : n1.z;
    auto proj = [&](const vec3 &v, int plane) {
        return plane == 0 ? p2{v.z, v.y} : (plane == 1 ? p2{v.x, v.z} : p2{v.x, v.y});
    };
    int plane = ((nx > nz) && (nx >= ny)) ? 0 : (((ny > nz) && (ny >= nx)) ? 1 : 2);
    if(plane == 2)
        return overlap_2d(proj(p1, 2), proj(q1, 2), proj(r1, 2), proj(p2_, 2), proj(q2, 2), proj(r2, 2));
    // the YZ and XZ projections swap the first two vertices of each triangle
    return overlap_2d(proj(q1, plane), proj(p1, plane), proj(r1, plane), proj(q2, plane), proj(p2_, plane),
                      proj(r2, plane));
}

//! the two interval checks once T1 = (p1, q1, r1) has p1 alone on its side and T2 is permuted to match
VK_HD bool check_min_max(const vec3 &p1, const vec3 &q1, const vec3 &r1, const vec3 &p2_, const vec3 &q2,
                         const vec3 &r2)
{
    vec3 n = cross(p2_ - q1, p1 - q1);
    if(dot(q2 - q1, n) >= -EPS) return false;
    n = cross(p2_ - p1, r1 - p1);
    if(dot(r2 - p1, n) >= -EPS) return false;
    return true;
}

//! canonical permutation of T2 by the signs of its distances to T1's plane
VK_HD bool permute_t2(const vec3 &p1, const vec3 &q1, const vec3 &r1, const vec3 &p2_, const vec3 &q2,
                      const vec3 &r2, float dp2, float dq2, float dr2, const vec3 &n1)
{
    if(dp2 > 0.f)
    {
        if(dq2 > 0.f) return check_min_max(p1, r1, q1, r2, p2_, q2);
        if(dr2 > 0.f) return check_min_max(p1, r1, q1, q2, r2, p2_);
        return check_min_max(p1, q1, r1, p2_, q2, r2);
    }
    if(dp2 < 0.f)
    {
        if(dq2 < 0.f) return check_min_max(p1, q1, r1, r2, p2_, q2);
        if(dr2 < 0.f) return check_min_max(p1, q1, r1, q2, r2, p2_);
        return check_min_max(p1, r1, q1, p2_, q2, r2);
    }
    if(dq2 < 0.f)
    {
        if(dr2 >= 0.f) return check_min_max(p1, r1, q1, q2, r2, p2_);
        return check_min_max(p1, q1, r1, p2_, q2, r2);
    }
    if(dq2 > 0.f)
    {
        if(dr2 > 0.f) return check_min_max(p1, r1, q1, p2_, q2, r2);
        return check_min_max(p1, q1, r1, q2, r2, p2_);
    }
    if(dr2 > 0.f) return check_min_max(p1, q1, r1, r2, p2_, q2);
    if(dr2 < 0.f) return check_min_max(p1, r1, q1, r2, p2_, q2);
    return coplanar(p1, q1, r1, p2_, q2, r2, n1);
}
}// namespace gd

VK_HD uint32_t tri_tri_overlap(const float a[9], const float b[9])
{
    using namespace gd;
    const vec3 p1 = ld(a), q1 = ld(a + 3), r1 = ld(a + 6), P2 = ld(b), Q2 = ld(b + 3), R2 = ld(b + 6);
    // signed distances of T1's vertices to T2's plane
    const vec3 n2 = cross(P2 - R2, Q2 - R2);
    float dp1 = dot(p1 - R2, n2), dq1 = dot(q1 - R2, n2), dr1 = dot(r1 - R2, n2);
    if(((dp1 * dq1) > 0.f) && ((dp1 * dr1) > 0.f)) return 0;
    // signed distances of T2's vertices to T1's plane
    const vec3 n1 = cross(q1 - p1, r1 - p1);
    float dp2 = dot(P2 - r1, n1), dq2 = dot(Q2 - r1, n1), dr2 = dot(R2 - r1, n1);
    if(((dp2 * dq2) > 0.f) && ((dp2 * dr2) > 0.f)) return 0;
    auto snap = [](float &d) {
        if(fabs_(d) <= EPS) d = 0.f;
    };
    snap(dp1), snap(dq1), snap(dr1), snap(dp2), snap(dq2), snap(dr2);
    bool r;
    if(dp1 > 0.f)
    {
        if(dq1 > 0.f) r = permute_t2(r1, p1, q1, P2, R2, Q2, dp2, dr2, dq2, n1);
        else if(dr1 > 0.f) r = permute_t2(q1, r1, p1, P2, R2, Q2, dp2, dr2, dq2, n1);
        else r = permute_t2(p1, q1, r1, P2, Q2, R2, dp2, dq2, dr2, n1);
    }
    else if(dp1 < 0.f)
    {
        if(dq1 < 0.f) r = permute_t2(r1, p1, q1, P2, Q2, R2, dp2, dq2, dr2, n1);
        else if(dr1 < 0.f) r = permute_t2(q1, r1, p1, P2, Q2, R2, dp2, dq2, dr2, n1);
        else r = permute_t2(p1, q1, r1, P2, R2, Q2, dp2, dr2, dq2, n1);
    }
    else if(dq1 < 0.f)
    {
        if(dr1 >= 0.f) r = permute_t2(q1, r1, p1, P2, R2, Q2, dp2, dr2, dq2, n1);
        else r = permute_t2(p1, q1, r1, P2, Q2, R2, dp2, dq2, dr2, n1);
    }
    else if(dq1 > 0.f)
    {
        if(dr1 > 0.f) r = permute_t2(p1, q1, r1, P2, R2, Q2, dp2, dr2, dq2, n1);
        else r = permute_t2(q1, r1, p1, P2, Q2, R2, dp2, dq2, dr2, n1);
    }
    else
    {
        if(dr1 > 0.f) r = permute_t2(r1, p1, q1, P2, Q2, R2, dp2, dq2, dr2, n1);
        else if(dr1 < 0.f) r = permute_t2(r1, p1, q1, P2, R2, Q2, dp2, dr2, dq2, n1);
        else r = coplanar(p1, q1, r1, P2, Q2, R2, n1);
    }
    return r ? 1u : 0u;
}

}// namespace vierkant::rt
