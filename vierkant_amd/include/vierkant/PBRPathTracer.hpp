// PBRPathTracer.hpp - the path tracer for C++ callers, with the reference's settings and accumulation
// semantics (include/vierkant/PBRPathTracer.hpp:42-142, src/PBRPathTracer.cpp:184-255, :677-682).
//
// settings_t keeps the reference's field names and defaults for everything that changes the traced image.
// image() follows the reference's display side: denoising selects the a-trous denoiser, and tonemap, gamma
// and exposure the composite of bloom_composition.slang. Bloom is not provided (the image is the reference's
// with bloom off); bloom and the other display-only fields are kept for source compatibility. Settings are
// copied at render_scene, as the reference copies them into its frame context. render_scene traces one batch
// unless max_num_batches is reached with suspend_trace_when_done set, then advances the batch index as
// min(batch + 1, max_num_batches).
#pragma once

#include <chrono>
#include <cmath>
#include <deque>
#include <memory>
#include <optional>
#include <vector>

#include "vierkant/RayBuilder.hpp"

namespace vierkant
{

//! a sun folded into the light list as its first light (include/vierkant/PBRPathTracer.hpp:23-33)
struct sunlight_params_t
{
    float color[3] = {0.f, 0.f, 0.f};
    float intensity = 0.f;
    //! elevation and azimuth in radians
    float spherical_coords[2] = {0.52359877559829887f, 0.34906585039886591f};
    //! apex half-angle of the sun disc in radians; 0 is a delta light
    float angular_size = 0.f;
};

//! light_t, the packed light record (include/vierkant/punctual_light.hpp:64-89)
using light_t = vkrt_light;

//! the sun's light record as the reference builds it (src/PBRPathTracer.cpp:611-624): propagation direction
//! quat(vec3(elevation, azimuth, 0)) * (0, 0, 1) = (cos e sin a, -sin e, cos e cos a), normalized
inline light_t sun_light(const sunlight_params_t &p)
{
    light_t l{};
    l.type = VKRT_LIGHT_DIRECTIONAL;
    for(int k = 0; k < 3; ++k) l.color[k] = p.color[k];
    l.intensity = p.intensity;
    const float e = p.spherical_coords[0], a = p.spherical_coords[1];
    float d[3] = {std::cos(e) * std::sin(a), -std::sin(e), std::cos(e) * std::cos(a)};
    const float len2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if(len2 > 0.f)
        for(float &c: d) c /= std::sqrt(len2);
    for(int k = 0; k < 3; ++k) l.direction[k] = d[k];
    l.range = INFINITY;
    l.angular_size = p.angular_size;
    return l;
}

class PBRPathTracer
{
public:
    struct settings_t
    {
        uint32_t resolution[2] = {1280, 720};
        uint32_t max_num_batches = 0;
        uint32_t num_samples = 1;
        uint32_t max_trace_depth = 6;
        float max_path_beta = 0.f;
        bool suspend_trace_when_done = true;
        bool disable_material = false;
        uint32_t mis_mode = 0;
        bool suppress_refractive_caustics = false;
        bool draw_skybox = true;
        bool compaction = true;
        bool denoising = false;
        bool tonemap = true;
        bool bloom = true;
        float gamma = 1.0f;
        float exposure = 2.0f;
        bool depth_of_field = false;
        bool suppress_reset = false;
        //! optional sun, folded in as the first light
        std::optional<sunlight_params_t> sunlight_params;
        uint32_t timing_history_size = 300;
    };

    struct create_info_t
    {
        uint32_t seed = 0;
        settings_t settings = {};
    };

    struct statistics_t
    {
        std::chrono::steady_clock::time_point timestamp;
        double raytrace_ms = 0.0;
        uint64_t segments = 0;
    };

    using scene_t = vkrt_pt_scene;
    using camera_t = vkrt_camera;

    static std::shared_ptr<PBRPathTracer> create(gpu_context_ptr ctx, const create_info_t &info)
    {
        return std::shared_ptr<PBRPathTracer>(new PBRPathTracer(std::move(ctx), info));
    }

    PBRPathTracer(const PBRPathTracer &) = delete;
    PBRPathTracer &operator=(const PBRPathTracer &) = delete;
    ~PBRPathTracer() { vkrt_pt_destroy(m_pt); }

    //! upload a scene and build its acceleration structure
    void set_scene(const scene_t &scene) { m_ctx->check(vkrt_pt_set_scene(m_pt, &scene)); }

    using instance_t = vkrt_pt_instance;

    //! an instanced scene: the scene's triangles split into meshes at mesh_first (n_meshes + 1 offsets),
    //! one BVH per mesh, placed by `instances` (the top level RayBuilder builds over its entries)
    void set_scene(const scene_t &scene, const std::vector<uint32_t> &mesh_first,
                   const std::vector<instance_t> &instances)
    {
        if(mesh_first.size() < 2) throw std::runtime_error("vkrt: mesh_first needs n_meshes + 1 offsets");
        m_ctx->check(vkrt_pt_set_instanced_scene(m_pt, &scene, mesh_first.data(),
                                                 static_cast<uint32_t>(mesh_first.size() - 1), instances.data(),
                                                 static_cast<uint32_t>(instances.size())));
    }

    //! the scene's light sources (the reference gathers them from lightsource components each frame)
    void set_lights(std::vector<light_t> lights) { m_lights = std::move(lights); }

    //! one frame: settings are copied and the light list (sun first) is gathered, a batch is traced unless
    //! suspended
    void render_scene(const camera_t &camera)
    {
        vkrt_pt_settings s = to_c(settings);
        m_ctx->check(vkrt_pt_set_settings(m_pt, &s));
        std::vector<light_t> frame_lights;
        frame_lights.reserve(m_lights.size() + 1);
        if(settings.sunlight_params && settings.sunlight_params->intensity > 0.f)
            frame_lights.push_back(sun_light(*settings.sunlight_params));
        frame_lights.insert(frame_lights.end(), m_lights.begin(), m_lights.end());
        m_ctx->check(vkrt_pt_set_lights(m_pt, frame_lights.data(), static_cast<uint32_t>(frame_lights.size())));
        m_ctx->check(vkrt_pt_render(m_pt, &camera));
        vkrt_pt_stats st{};
        vkrt_pt_get_stats(m_pt, &st);
        if(st.traced)
        {
            m_statistics.push_back({std::chrono::steady_clock::now(), st.raytrace_ms, st.segments});
            while(m_statistics.size() > settings.timing_history_size) m_statistics.pop_front();
        }
    }

    [[nodiscard]] size_t current_batch() const { return vkrt_pt_current_batch(m_pt); }

    void reset_accumulator()
    {
        vkrt_pt_settings s = to_c(settings);
        m_ctx->check(vkrt_pt_set_settings(m_pt, &s));
        m_ctx->check(vkrt_pt_reset_accumulator(m_pt));
    }

    //! the pixel_data_t buffer, row-major from the top-left
    [[nodiscard]] std::vector<vkrt_pixel> pixels() const
    {
        vkrt_pt_settings s{};
        vkrt_pt_get_settings(m_pt, &s);
        std::vector<vkrt_pixel> out((size_t)s.resolution[0] * s.resolution[1]);
        m_ctx->check(vkrt_pt_read_pixels(m_pt, out.data(), 0));
        return out;
    }

    //! the output image (float4 per pixel) with the settings of the last render_scene: denoised when
    //! denoising (PBRPathTracer::denoise_pass), then composited when tonemap (PBRPathTracer::post_fx_pass)
    [[nodiscard]] std::vector<float> image() const
    {
        vkrt_pt_settings s{};
        vkrt_pt_get_settings(m_pt, &s);
        std::vector<float> out(4 * (size_t)s.resolution[0] * s.resolution[1]);
        m_ctx->check(vkrt_pt_read_image(m_pt, out.data(), 0));
        return out;
    }

    [[nodiscard]] const std::deque<statistics_t> &statistics() const { return m_statistics; }

    settings_t settings;

private:
    PBRPathTracer(gpu_context_ptr ctx, const create_info_t &info) : settings(info.settings), m_ctx(std::move(ctx))
    {
        vkrt_pt_settings s = to_c(settings);
        m_ctx->check(vkrt_pt_create(m_ctx->handle(), &s, info.seed, &m_pt));
    }

    static vkrt_pt_settings to_c(const settings_t &s)
    {
        vkrt_pt_settings c{};
        c.resolution[0] = s.resolution[0];
        c.resolution[1] = s.resolution[1];
        c.max_num_batches = s.max_num_batches;
        c.num_samples = s.num_samples;
        c.max_trace_depth = s.max_trace_depth;
        c.max_path_beta = s.max_path_beta;
        c.suspend_trace_when_done = s.suspend_trace_when_done ? 1u : 0u;
        c.disable_material = s.disable_material ? 1u : 0u;
        c.suppress_reset = s.suppress_reset ? 1u : 0u;
        c.denoising = s.denoising ? 1u : 0u;
        c.tonemap = s.tonemap ? 1u : 0u;
        c.mis_mode = s.mis_mode;
        c.suppress_refractive_caustics = s.suppress_refractive_caustics ? 1u : 0u;
        c.gamma = s.gamma;
        c.exposure = s.exposure;
        return c;
    }

    gpu_context_ptr m_ctx;
    vkrt_pt *m_pt = nullptr;
    std::vector<light_t> m_lights;
    std::deque<statistics_t> m_statistics;
};

}// namespace vierkant
