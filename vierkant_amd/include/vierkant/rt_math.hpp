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
VK_HD bool mt_fast(const vec3 &o, const vec3 &d, const vec3 &v0, const vec3 &e1, const vec3 &e2, float &t, float &u,
                   float &v, bool cull_back = false)
{
    vec3 pvec = cross(d, e2);
    float det = dot(e1, pvec);
    constexpr float epsilon = 10.0e-10f;
    if(det > -epsilon && det < epsilon) return false;
    if(cull_back && det < 0.f) return false;
    vec3 tvec = o - v0;
    float a = dot(tvec, pvec);
    const bool neg = det < 0.f;
    const float D = neg ? -det : det;
    const bool bounded = D <= 0x1p100f;
    const float lo = D * 0x1p-80f;
    const float A = neg ? -a : a;
    if(bounded && (A <= -lo || A >= D * (1.0f + 0x1p-19f))) return false;
    vec3 qvec = cross(tvec, e1);
    float b = dot(d, qvec);
    const float B = neg ? -b : b;
    if(bounded && (B <= -lo || (A + B) >= D * (1.0f + 0x1p-16f))) return false;
    // reference tail (src/intersection.cpp:121-128)
    float inv_det = 1.0f / det;
    float uu = inv_det * a;
    if(uu < 0.0f || uu > 1.0f) return false;
    float vv = b * inv_det;
    if(vv < 0.0f || (uu + vv) > 1.0f) return false;
    t = dot(e2, qvec) * inv_det;
    u = uu;
    v = vv;
    return true;
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

}// namespace vierkant::rt
