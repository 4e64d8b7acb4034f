// RayBuilder.hpp - mesh -> acceleration structure and ray queries on the MI355X core, for C++ callers.
//
// Mirrors the role of vierkant::RayBuilder (include/vierkant/RayBuilder.hpp:21-336 of the reference):
// build an acceleration structure per mesh entry (create_mesh_structures, src/RayBuilder.cpp:98-303),
// refit it when vertices move (the update path of build_scene_acceleration, :783-1035) and report build
// timings (timings(), :1079-1108). Everything goes through the C ABI in vkrt.h; errors become
// std::runtime_error as in the reference (src/RayBuilder.cpp:174).
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "vkrt.h"

namespace vierkant
{

//! an owned vkrt context (device + HIP stream)
class gpu_context
{
public:
    explicit gpu_context(int device = 0, void *hip_stream = nullptr)
    {
        int rc = vkrt_create(device, hip_stream, &m_ctx);
        if(rc != VKRT_OK) throw std::runtime_error("vkrt_create failed: status " + std::to_string(rc));
    }
    gpu_context(const gpu_context &) = delete;
    gpu_context &operator=(const gpu_context &) = delete;
    ~gpu_context() { vkrt_destroy(m_ctx); }

    [[nodiscard]] vkrt_ctx *handle() const { return m_ctx; }

    void check(int rc) const
    {
        if(rc != VKRT_OK) throw std::runtime_error(std::string("vkrt: ") + vkrt_last_error(m_ctx));
    }

private:
    vkrt_ctx *m_ctx = nullptr;
};

using gpu_context_ptr = std::shared_ptr<gpu_context>;

class RayBuilder
{
public:
    //! one mesh entry: positions with a byte stride, 3 indices per triangle (vertex_offset/base_index applied)
    struct mesh_t
    {
        const float *positions = nullptr;
        uint32_t num_vertices = 0;
        uint32_t stride = 12;
        const uint32_t *indices = nullptr;
        uint32_t num_indices = 0;
        bool device_pointers = false;
    };

    struct build_params_t
    {
        uint32_t builder = VKRT_BUILDER_PLOC;
        uint32_t max_leaf_size = 2;
        uint32_t width = 0;// 0: 4-wide traversal layout when it fits, 2: binary tree only
    };

    struct timings_t
    {
        double build_ms = 0.0;
    };

    //! an acceleration structure (BLAS of one mesh entry)
    class acceleration_structure
    {
    public:
        explicit acceleration_structure(vkrt_bvh *h) : m_bvh(h) {}
        acceleration_structure(const acceleration_structure &) = delete;
        acceleration_structure &operator=(const acceleration_structure &) = delete;
        ~acceleration_structure() { vkrt_bvh_destroy(m_bvh); }
        [[nodiscard]] vkrt_bvh *handle() const { return m_bvh; }
        [[nodiscard]] vkrt_bvh_info info() const
        {
            vkrt_bvh_info i{};
            vkrt_bvh_get_info(m_bvh, &i);
            return i;
        }

    private:
        vkrt_bvh *m_bvh = nullptr;
    };
    using acceleration_structure_ptr = std::unique_ptr<acceleration_structure>;

    explicit RayBuilder(gpu_context_ptr ctx) : m_ctx(std::move(ctx)) {}

    acceleration_structure_ptr create_mesh_structure(const mesh_t &mesh)
    {
        return create_mesh_structure(mesh, build_params_t{});
    }

    acceleration_structure_ptr create_mesh_structure(const mesh_t &mesh, const build_params_t &params)
    {
        vkrt_bvh_params p{};
        p.builder = params.builder;
        p.max_leaf_size = params.max_leaf_size;
        p.width = params.width;
        vkrt_bvh *h = nullptr;
        m_ctx->check(vkrt_bvh_build(m_ctx->handle(), mesh.positions, mesh.num_vertices, mesh.stride, mesh.indices,
                                    mesh.num_indices / 3, &p, mesh.device_pointers ? VKRT_DEVICE_PTRS : 0, &h));
        auto as = std::make_unique<acceleration_structure>(h);
        m_timings.build_ms = as->info().build_ms;
        return as;
    }

    //! new vertex positions, same topology
    void refit(acceleration_structure &as, const float *positions, uint32_t stride = 12, bool device_pointers = false)
    {
        m_ctx->check(vkrt_bvh_refit(m_ctx->handle(), as.handle(), positions, stride,
                                    device_pointers ? VKRT_DEVICE_PTRS : 0));
        m_timings.build_ms = as.info().build_ms;
    }

    //! closest hits (prim = VKRT_INVALID_PRIM on a miss), identical to the brute-force reference loop
    std::vector<vkrt_hit> trace_closest(const acceleration_structure &as, const std::vector<vkrt_ray> &rays) const
    {
        std::vector<vkrt_hit> hits(rays.size());
        m_ctx->check(vkrt_trace_closest(m_ctx->handle(), as.handle(), rays.data(), static_cast<uint32_t>(rays.size()),
                                        hits.data(), 0));
        return hits;
    }

    //! shadow queries: 1 if anything is hit inside (tmin, tmax)
    std::vector<uint8_t> trace_any(const acceleration_structure &as, const std::vector<vkrt_ray> &rays) const
    {
        std::vector<uint8_t> occ(rays.size());
        m_ctx->check(vkrt_trace_any(m_ctx->handle(), as.handle(), rays.data(), static_cast<uint32_t>(rays.size()),
                                    occ.data(), 0));
        return occ;
    }

    //! a two-level scene: instances (transform + mesh index) of acceleration structures, the role of the
    //! reference's top-level structure (src/RayBuilder.cpp:386-781)
    class scene
    {
    public:
        scene(gpu_context_ptr ctx, const std::vector<const acceleration_structure *> &meshes,
              const std::vector<vkrt_instance> &instances)
            : m_ctx(std::move(ctx))
        {
            std::vector<vkrt_bvh *> h;
            for(const auto *m: meshes) h.push_back(m->handle());
            m_ctx->check(vkrt_scene_create(m_ctx->handle(), h.data(), static_cast<uint32_t>(h.size()), instances.data(),
                                           static_cast<uint32_t>(instances.size()), &m_scene));
        }
        scene(const scene &) = delete;
        scene &operator=(const scene &) = delete;
        ~scene() { vkrt_scene_destroy(m_scene); }

        //! new transforms or refitted meshes: the top level is rebuilt
        void update(const std::vector<vkrt_instance> &instances)
        {
            m_ctx->check(vkrt_scene_set_instances(m_scene, instances.data(), static_cast<uint32_t>(instances.size())));
        }

        std::vector<vkrt_instance_hit> trace_closest(const std::vector<vkrt_ray> &rays) const
        {
            std::vector<vkrt_instance_hit> hits(rays.size());
            m_ctx->check(vkrt_scene_trace_closest(m_ctx->handle(), m_scene, rays.data(),
                                                  static_cast<uint32_t>(rays.size()), hits.data(), 0));
            return hits;
        }

    private:
        gpu_context_ptr m_ctx;
        vkrt_scene *m_scene = nullptr;
    };

    //! an instance of mesh `mesh` with the given 3x4 row-major object-to-world matrix
    static vkrt_instance make_instance(const float object_to_world[12], uint32_t mesh, uint32_t custom_index = 0)
    {
        vkrt_instance in{};
        vkrt_instance_set_transform(&in, object_to_world);
        in.blas = mesh;
        in.custom_index = custom_index;
        return in;
    }

    [[nodiscard]] timings_t timings() const { return m_timings; }

private:
    gpu_context_ptr m_ctx;
    timings_t m_timings;
};

}// namespace vierkant
