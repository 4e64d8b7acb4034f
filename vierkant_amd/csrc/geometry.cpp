// geometry.cpp - procedural meshes on the host: Geometry::IcoSphere and Geometry::Plane.
//
// Vertex/index order and arithmetic follow src/Geometry.cpp:566-613 (IcoSphere), :141-192
// (tessellate: three fresh vertex copies per triangle and level, index pattern
// {a, k+1, k, k+1, b, k+2, k, k+2, c, k, k+1, k+2}) and :313-371 (Plane in the XZ plane), so that
// the scenes built from them have the reference's triangle order and vertex bits.
#include <cmath>
#include <cstring>
#include <vector>

#include "context.hpp"
#include "vierkant/rt_math.hpp"

using vierkant::rt::vec3;

namespace vkrt
{

static vec3 normalize3(const vec3 &v)
{
    float inv = 1.0f / std::sqrt(vierkant::rt::dot(v, v));
    return v * inv;
}

void icosphere_counts(uint32_t subdiv, uint32_t &n_vert, uint32_t &n_idx)
{
    uint64_t nv = 12, nt = 20;
    for(uint32_t l = 0; l < subdiv; ++l)
    {
        nv += 3 * nt;
        nt *= 4;
    }
    n_vert = static_cast<uint32_t>(nv);
    n_idx = static_cast<uint32_t>(3 * nt);
}

void icosphere(float radius, uint32_t subdiv, std::vector<vec3> &verts, std::vector<uint32_t> &indices)
{
    constexpr float phi = 1.61803398874989484820f;
    const float e1 = radius / std::sqrt(1.f + phi * phi);
    const float e2 = phi * e1;
    verts = {{-e1, e2, 0}, {e1, e2, 0},   {-e1, -e2, 0}, {e1, -e2, 0}, {0, -e1, e2},  {0, e1, e2},
             {0, -e1, -e2}, {0, e1, -e2}, {e2, 0, -e1},  {e2, 0, e1},  {-e2, 0, -e1}, {-e2, 0, e1}};
    indices = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4,  11, 10, 2,  10, 7, 6, 7, 1, 8,
               3, 9,  4, 3, 4, 2, 3, 2, 6, 3, 6, 8,  3, 8,  9,  4, 9, 5, 2, 4,  11, 6,  2,  10, 8,  6, 7, 9, 8, 1};
    uint32_t nv_total, ni_total;
    icosphere_counts(subdiv, nv_total, ni_total);
    verts.reserve(nv_total);
    std::vector<uint32_t> next;
    for(uint32_t level = 0; level < subdiv; ++level)
    {
        next.clear();
        next.reserve(indices.size() * 4);
        for(size_t j = 0; j < indices.size(); j += 3)
        {
            const uint32_t k = static_cast<uint32_t>(verts.size());
            const uint32_t a = indices[j], b = indices[j + 1], c = indices[j + 2];
            const uint32_t quad[12] = {a, k + 1, k, k + 1, b, k + 2, k, k + 2, c, k, k + 1, k + 2};
            next.insert(next.end(), quad, quad + 12);
            const vec3 va = verts[a], vb = verts[b], vc = verts[c];
            // slot order: k = ac, k + 1 = ab, k + 2 = bc
            verts.push_back(radius * normalize3(va + vc));
            verts.push_back(radius * normalize3(va + vb));
            verts.push_back(radius * normalize3(vb + vc));
        }
        indices.swap(next);
    }
}

void plane(float width, float height, uint32_t sw, uint32_t sh, std::vector<vec3> &verts,
           std::vector<uint32_t> &indices)
{
    const float width_half = width / 2, height_half = height / 2;
    const float segment_width = width / sw, segment_height = height / sh;
    const uint32_t gx1 = sw + 1, gz1 = sh + 1;
    verts.clear();
    verts.reserve((size_t)gx1 * gz1);
    for(uint32_t iz = 0; iz < gz1; ++iz)
        for(uint32_t ix = 0; ix < gx1; ++ix)
        {
            float x = ix * segment_width - width_half;
            float z = iz * segment_height - height_half;
            verts.push_back({x, 0.f, z});
        }
    indices.clear();
    indices.reserve((size_t)sw * sh * 6);
    for(uint32_t iz = 0; iz < sh; ++iz)
        for(uint32_t ix = 0; ix < sw; ++ix)
        {
            uint32_t a = ix + gx1 * iz, b = ix + gx1 * (iz + 1);
            uint32_t c = (ix + 1) + gx1 * (iz + 1), d = (ix + 1) + gx1 * iz;
            const uint32_t two[6] = {a, b, c, c, d, a};
            indices.insert(indices.end(), two, two + 6);
        }
}

}// namespace vkrt

extern "C" {

int vkrt_icosphere_size(uint32_t subdiv, uint32_t *n_vert, uint32_t *n_idx)
{
    if(!n_vert || !n_idx || subdiv > 10) return VKRT_E_INVALID_ARG;
    vkrt::icosphere_counts(subdiv, *n_vert, *n_idx);
    return VKRT_OK;
}

int vkrt_icosphere(float radius, uint32_t subdiv, float *pos, uint32_t *idx)
{
    if(!pos || !idx || subdiv > 10) return VKRT_E_INVALID_ARG;
    try
    {
        std::vector<vec3> v;
        std::vector<uint32_t> i;
        vkrt::icosphere(radius, subdiv, v, i);
        std::memcpy(pos, v.data(), v.size() * sizeof(vec3));
        std::memcpy(idx, i.data(), i.size() * sizeof(uint32_t));
    } catch(...)
    {
        return VKRT_E_OOM;
    }
    return VKRT_OK;
}

int vkrt_plane_size(uint32_t sw, uint32_t sh, uint32_t *n_vert, uint32_t *n_idx)
{
    if(!n_vert || !n_idx || sw == 0 || sh == 0) return VKRT_E_INVALID_ARG;
    uint64_t nv = (uint64_t)(sw + 1) * (sh + 1), ni = (uint64_t)sw * sh * 6;
    if(nv > 0xffffffffu || ni > 0xffffffffu) return VKRT_E_INVALID_ARG;
    *n_vert = (uint32_t)nv;
    *n_idx = (uint32_t)ni;
    return VKRT_OK;
}

int vkrt_plane(float width, float height, uint32_t sw, uint32_t sh, float *pos, uint32_t *idx)
{
    if(!pos || !idx || sw == 0 || sh == 0) return VKRT_E_INVALID_ARG;
    try
    {
        std::vector<vec3> v;
        std::vector<uint32_t> i;
        vkrt::plane(width, height, sw, sh, v, i);
        std::memcpy(pos, v.data(), v.size() * sizeof(vec3));
        std::memcpy(idx, i.data(), i.size() * sizeof(uint32_t));
    } catch(...)
    {
        return VKRT_E_OOM;
    }
    return VKRT_OK;
}

}// extern "C"
