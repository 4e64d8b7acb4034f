// intersection.hpp - vierkant's intersection API (include/vierkant/intersection.hpp of the reference) for
// C++ callers of the MI355X core: the same types and intersect(...) overloads with the reference's results.
//
// The scalar overloads evaluate the reference arithmetic of vierkant/rt_math.hpp on the calling thread,
// exactly as the reference's free functions do; batched work (many rays, many primitives) goes to the GPU
// through the C ABI (vkrt.h) via vierkant::RayBuilder and the vkrt_intersect_* entry points.
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

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

// ---- point sets (src/intersection.cpp:25-110) -----------------------------------------------------

//! compute_aabb (src/intersection.cpp:96-106): glm::min / glm::max fold from the empty box
inline AABB compute_aabb(const std::vector<vec3> &vertices)
{
    AABB ret;
    for(const vec3 &p: vertices)
    {
        ret.min = vec3{p.x < ret.min.x ? p.x : ret.min.x, p.y < ret.min.y ? p.y : ret.min.y, p.z < ret.min.z ? p.z : ret.min.z};
        ret.max = vec3{ret.max.x < p.x ? p.x : ret.max.x, ret.max.y < p.y ? p.y : ret.max.y, ret.max.z < p.z ? p.z : ret.max.z};
    }
    return ret;
}

//! compute_bounding_sphere (src/intersection.cpp:25-92): Ritter's sphere; vkrt_compute_bounding_sphere is
//! the same computation on the GPU. An empty array gives a zero sphere (the reference reads vertices[0]).
inline Sphere compute_bounding_sphere(const std::vector<vec3> &vertices)
{
    Sphere ret;
    if(vertices.empty()) return ret;
    auto at = [&](size_t i, int a) { return a == 0 ? vertices[i].x : (a == 1 ? vertices[i].y : vertices[i].z); };
    size_t pmin[3] = {0, 0, 0}, pmax[3] = {0, 0, 0};
    for(size_t i = 0; i < vertices.size(); ++i)
        for(int a = 0; a < 3; ++a)
        {
            pmin[a] = at(i, a) < at(pmin[a], a) ? i : pmin[a];
            pmax[a] = at(i, a) > at(pmax[a], a) ? i : pmax[a];
        }
    float best = 0.f;
    int axis = 0;
    for(int a = 0; a < 3; ++a)
    {
        const vec3 d = vertices[pmax[a]] - vertices[pmin[a]];
        const float d2 = rt::dot(d, d);
        if(d2 > best)
        {
            best = d2;
            axis = a;
        }
    }
    const vec3 &p1 = vertices[pmin[axis]], &p2 = vertices[pmax[axis]];
    ret.center = vec3{(p1.x + p2.x) / 2.f, (p1.y + p2.y) / 2.f, (p1.z + p2.z) / 2.f};
    ret.radius = std::sqrt(best) / 2.f;
    for(const vec3 &p: vertices)
    {
        const vec3 v = p - ret.center;
        const float d2 = rt::dot(v, v);
        if(d2 > ret.radius * ret.radius)
        {
            const float d = std::sqrt(d2);
            const float k = 0.5f + (ret.radius / d) / 2;
            ret.center = vec3{p.x * (1.f - k) + ret.center.x * k, p.y * (1.f - k) + ret.center.y * k,
                              p.z * (1.f - k) + ret.center.z * k};
            ret.radius = (ret.radius + d) / 2;
        }
    }
    return ret;
}

//! compute_centroid (src/intersection.cpp:110) returns crocore::mean, whose source is not in the reference
//! checkout (empty submodule): here the sequential mean, sum in index order divided by the count
inline vec3 compute_centroid(const std::vector<vec3> &vertices)
{
    vec3 s{0.f, 0.f, 0.f};
    for(const vec3 &p: vertices) s = vec3{s.x + p.x, s.y + p.y, s.z + p.z};
    const float n = static_cast<float>(vertices.size());
    return vertices.empty() ? s : vec3{s.x / n, s.y / n, s.z / n};
}

}// namespace vierkant
