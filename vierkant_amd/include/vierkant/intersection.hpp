// intersection.hpp - vierkant's intersection API (include/vierkant/intersection.hpp of the reference) for
// C++ callers of the MI355X core: the same types and intersect(...) overloads with the reference's results.
//
// The scalar overloads evaluate the reference arithmetic of vierkant/rt_math.hpp on the calling thread,
// exactly as the reference's free functions do; batched work (many rays, many primitives) goes to the GPU
// through the C ABI (vkrt.h) via vierkant::RayBuilder and the vkrt_intersect_* entry points.
#pragma once

#include <cstdint>

#include "vierkant/rt_math.hpp"

namespace vierkant
{

enum intersection_type
{
    REJECT = 0,
    INTERSECT = 1,
    INSIDE = 2
};

using vec3 = rt::vec3;

//! type of intersection and distance along a ray (intersection.hpp:28-38)
struct ray_intersection
{
    intersection_type type = REJECT;
    float distance = 0.f;

    ray_intersection(intersection_type t = REJECT, float d = 0.f) : type(t), distance(d) {}
    explicit operator bool() const { return type != REJECT; }
};

//! plus barycentric coordinates within the triangle (intersection.hpp:44-51)
struct ray_triangle_intersection : public ray_intersection
{
    float u = 0.f, v = 0.f;

    ray_triangle_intersection(intersection_type t = REJECT, float d = 0.f, float u_ = 0.f, float v_ = 0.f)
        : ray_intersection(t, d), u(u_), v(v_)
    {}
};

//! origin + direction; the constructor normalizes the direction with glm's order (intersection.hpp:145-152)
struct Ray
{
    vec3 origin{0.f, 0.f, 0.f};
    vec3 direction{0.f, 0.f, 1.f};

    Ray() = default;
    Ray(const vec3 &o, const vec3 &d) : origin(o)
    {
        float inv = 1.0f / std::sqrt(rt::dot(d, d));
        direction = d * inv;
    }
};

struct Triangle
{
    vec3 v0, v1, v2;
};

struct Sphere
{
    vec3 center{0.f, 0.f, 0.f};
    float radius = 0.f;
};

//! Ax + By + Cz + D = 0 with a unit normal (A, B, C)
struct Plane
{
    float coefficients[4] = {0.f, 1.f, 0.f, 0.f};
};

struct AABB
{
    vec3 min{3.40282347e+38f, 3.40282347e+38f, 3.40282347e+38f};
    vec3 max{-3.40282347e+38f, -3.40282347e+38f, -3.40282347e+38f};

    AABB() = default;
    AABB(const vec3 &mn, const vec3 &mx) : min(mn), max(mx) {}
    [[nodiscard]] vec3 center() const { return vec3{(max.x + min.x) / 2.f, (max.y + min.y) / 2.f, (max.z + min.z) / 2.f}; }
    [[nodiscard]] vec3 half_extents() const { return vec3{(max.x - min.x) / 2.f, (max.y - min.y) / 2.f, (max.z - min.z) / 2.f}; }
    [[nodiscard]] bool valid() const { return max.x >= min.x && max.y >= min.y && max.z >= min.z; }
    AABB &operator+=(const AABB &o)
    {
        min = vec3{o.min.x < min.x ? o.min.x : min.x, o.min.y < min.y ? o.min.y : min.y, o.min.z < min.z ? o.min.z : min.z};
        max = vec3{max.x < o.max.x ? o.max.x : max.x, max.y < o.max.y ? o.max.y : max.y, max.z < o.max.z ? o.max.z : max.z};
        return *this;
    }
    //! AABB::intersect(const Ray&) (src/intersection.cpp:308-312)
    [[nodiscard]] ray_intersection intersect(const Ray &ray) const
    {
        float d = 0.f;
        uint32_t t = rt::aabb_reference(min, max, ray.origin, ray.direction, d);
        return t ? ray_intersection(INTERSECT, d) : ray_intersection(REJECT);
    }
};

//! oriented box: center, axes (columns) and half lengths (intersection.hpp:352-377)
struct OBB
{
    vec3 center{0.f, 0.f, 0.f};
    vec3 axis[3] = {{1.f, 0.f, 0.f}, {0.f, 1.f, 0.f}, {0.f, 0.f, 1.f}};
    vec3 half_lengths{0.f, 0.f, 0.f};
};

// ---- ray tests (src/intersection.cpp) ---------------------------------------------------------------

inline ray_triangle_intersection intersect(const Triangle &tri, const Ray &ray)
{
    float t = 0.f, u = 0.f, v = 0.f;
    if(rt::mt_reference(ray.origin, ray.direction, tri.v0, tri.v1 - tri.v0, tri.v2 - tri.v0, t, u, v))
        return ray_triangle_intersection(INTERSECT, t, u, v);
    return ray_triangle_intersection(REJECT);
}
inline ray_triangle_intersection intersect(const Ray &ray, const Triangle &tri) { return intersect(tri, ray); }

inline ray_intersection intersect(const OBB &obb, const Ray &ray)
{
    float d = 0.f;
    uint32_t t = rt::obb_reference(obb.center, obb.axis, obb.half_lengths, ray.origin, ray.direction, d);
    return t ? ray_intersection(INTERSECT, d) : ray_intersection(REJECT);
}
inline ray_intersection intersect(const Ray &ray, const OBB &obb) { return intersect(obb, ray); }

//! declared in the reference (intersection.hpp:90) without a definition; here it is AABB::intersect
inline ray_intersection intersect(const AABB &aabb, const Ray &ray) { return aabb.intersect(ray); }
inline ray_intersection intersect(const Ray &ray, const AABB &aabb) { return aabb.intersect(ray); }

inline ray_intersection intersect(const Sphere &s, const Ray &ray)
{
    float d = 0.f;
    uint32_t t = rt::sphere_reference(s.center, s.radius, ray.origin, ray.direction, d);
    return t ? ray_intersection(INTERSECT, d) : ray_intersection(REJECT);
}
inline ray_intersection intersect(const Ray &ray, const Sphere &s) { return intersect(s, ray); }

inline ray_intersection intersect(const Plane &p, const Ray &ray)
{
    float d = 0.f;
    uint32_t t = rt::plane_reference(p.coefficients, ray.origin, ray.direction, d);
    return t ? ray_intersection(INTERSECT, d) : ray_intersection(REJECT);
}
inline ray_intersection intersect(const Ray &ray, const Plane &p) { return intersect(p, ray); }

// ---- triangle tests (src/intersection.cpp:364-378) ----------------------------------------------------

inline uint32_t intersect(const Triangle &t, const AABB &b)
{
    const float tri[9] = {t.v0.x, t.v0.y, t.v0.z, t.v1.x, t.v1.y, t.v1.z, t.v2.x, t.v2.y, t.v2.z};
    const float mn[3] = {b.min.x, b.min.y, b.min.z}, mx[3] = {b.max.x, b.max.y, b.max.z};
    return rt::tri_box_overlap(tri, mn, mx);
}
inline uint32_t intersect(const AABB &b, const Triangle &t) { return intersect(t, b); }

inline uint32_t intersect(const Triangle &a, const Triangle &b)
{
    const float ta[9] = {a.v0.x, a.v0.y, a.v0.z, a.v1.x, a.v1.y, a.v1.z, a.v2.x, a.v2.y, a.v2.z};
    const float tb[9] = {b.v0.x, b.v0.y, b.v0.z, b.v1.x, b.v1.y, b.v1.z, b.v2.x, b.v2.y, b.v2.z};
    return rt::tri_tri_overlap(ta, tb);
}

}// namespace vierkant
