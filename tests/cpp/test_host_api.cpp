// test_host_api.cpp - the C++ host layer end to end on a GPU: RayBuilder's BVH queries equal the scalar
// vierkant::intersect(Triangle, Ray) brute-force loop bit for bit, and PBRPathTracer keeps the reference's
// batch semantics. Prints "host api ok" and exits 0 on success.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <vector>

#include "vierkant/PBRPathTracer.hpp"
#include "vierkant/RayBuilder.hpp"
#include "vierkant/intersection.hpp"
#include "vierkant/model.hpp"

using namespace vierkant;

static int fail(const char *what)
{
    std::printf("FAIL: %s\n", what);
    return 1;
}

static uint32_t bits(float f)
{
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return b;
}

int main()
{
    auto ctx = std::make_shared<gpu_context>(0);

    // a subdivision-3 icosphere from the library's Geometry::IcoSphere
    uint32_t nv = 0, ni = 0;
    vkrt_icosphere_size(3, &nv, &ni);
    std::vector<float> pos(3 * (size_t)nv);
    std::vector<uint32_t> idx(ni);
    vkrt_icosphere(1.0f, 3, pos.data(), idx.data());

    RayBuilder builder(ctx);
    RayBuilder::mesh_t mesh;
    mesh.positions = pos.data();
    mesh.num_vertices = nv;
    mesh.indices = idx.data();
    mesh.num_indices = ni;
    auto as = builder.create_mesh_structure(mesh);
    if(as->info().n_tri != ni / 3) return fail("triangle count");

    // deterministic rays (LCG), normalized by the Ray constructor
    uint32_t s = 12345u;
    auto rnd = [&]() {
        s = 1664525u * s + 1013904223u;
        return (float)(s >> 8) / 16777216.0f;
    };
    std::vector<vkrt_ray> rays;
    std::vector<Ray> host_rays;
    for(int i = 0; i < 4096; ++i)
    {
        vec3 o{rnd() * 6.f - 3.f, rnd() * 6.f - 3.f, rnd() * 6.f - 3.f};
        vec3 d{rnd() * 2.f - 1.f, rnd() * 2.f - 1.f, rnd() * 2.f - 1.f};
        Ray r(o, d);
        host_rays.push_back(r);
        rays.push_back(vkrt_ray{r.origin.x, r.origin.y, r.origin.z, 0.f, r.direction.x, r.direction.y,
                                r.direction.z, INFINITY});
    }
    auto hits = builder.trace_closest(*as, rays);
    auto occ = builder.trace_any(*as, rays);
    int n_hit = 0;
    for(size_t i = 0; i < rays.size(); ++i)
    {
        float best = INFINITY;
        uint32_t best_prim = VKRT_INVALID_PRIM;
        ray_triangle_intersection best_hit;
        for(uint32_t t = 0; t < ni / 3; ++t)
        {
            Triangle tri{{pos[3 * idx[3 * t]], pos[3 * idx[3 * t] + 1], pos[3 * idx[3 * t] + 2]},
                         {pos[3 * idx[3 * t + 1]], pos[3 * idx[3 * t + 1] + 1], pos[3 * idx[3 * t + 1] + 2]},
                         {pos[3 * idx[3 * t + 2]], pos[3 * idx[3 * t + 2] + 1], pos[3 * idx[3 * t + 2] + 2]}};
            auto h = intersect(host_rays[i], tri);
            if(h.type == INTERSECT && h.distance > 0.f && h.distance < best)
            {
                best = h.distance;
                best_prim = t;
                best_hit = h;
            }
        }
        if(hits[i].prim != best_prim) return fail("closest primitive differs from the scalar intersect loop");
        if(best_prim != VKRT_INVALID_PRIM)
        {
            ++n_hit;
            if(bits(hits[i].t) != bits(best_hit.distance) || bits(hits[i].u) != bits(best_hit.u) ||
               bits(hits[i].v) != bits(best_hit.v))
                return fail("t/u/v bits differ from the scalar intersect");
        }
        if(occ[i] != (best_prim != VKRT_INVALID_PRIM ? 1 : 0)) return fail("any-hit flag");
    }
    if(n_hit < 100) return fail("too few hits");

    // a two-level scene: two translated copies of the icosphere; every hit equals the plain BVH's hit of the
    // ray moved into the instance's space (translations by integers are exact)
    {
        const float m0[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
        const float m1[12] = {1, 0, 0, 4, 0, 1, 0, 0, 0, 0, 1, 0};
        std::vector<vkrt_instance> inst{RayBuilder::make_instance(m0, 0, 7), RayBuilder::make_instance(m1, 0, 8)};
        RayBuilder::scene sc(ctx, {as.get()}, inst);
        std::vector<vkrt_ray> r2 = rays;
        auto sh = sc.trace_closest(r2);
        int n_inst_hits = 0;
        for(size_t i = 0; i < rays.size(); ++i)
        {
            if(sh[i].prim == VKRT_INVALID_PRIM) continue;
            if(sh[i].instance == 0 && (sh[i].prim != hits[i].prim || bits(sh[i].t) != bits(hits[i].t)))
                return fail("scene instance 0 disagrees with the mesh BVH");
            n_inst_hits += sh[i].custom_index == 7 + sh[i].instance;
        }
        if(n_inst_hits < 100) return fail("scene hits");
        sc.update(inst);
    }

    // scalar volume tests of the API surface
    AABB box({0.f, 0.f, 0.f}, {1.f, 1.f, 1.f});
    if(box.intersect(Ray({-1.f, 0.5f, 0.5f}, {1.f, 0.f, 0.f})).distance != 1.0f) return fail("AABB entry distance");
    if(intersect(Sphere{{0.f, 0.f, 5.f}, 1.f}, Ray({0.f, 0.f, 0.f}, {0.f, 0.f, 1.f})).distance != 4.0f)
        return fail("sphere distance");

    // point sets: the scalar compute_bounding_sphere equals the GPU entry point bit for bit
    std::vector<vec3> cloud;
    for(uint32_t i = 0; i < nv; ++i) cloud.push_back(vec3{pos[3 * i] * 3.f, pos[3 * i + 1], pos[3 * i + 2] * 0.5f});
    const Sphere hs = compute_bounding_sphere(cloud);
    float gs[4];
    ctx->check(vkrt_compute_bounding_sphere(ctx->handle(), pos.data(), 0, gs, 0));
    std::vector<float> flat;
    for(const vec3 &p: cloud) flat.insert(flat.end(), {p.x, p.y, p.z});
    ctx->check(vkrt_compute_bounding_sphere(ctx->handle(), flat.data(), nv, gs, 0));
    if(bits(gs[0]) != bits(hs.center.x) || bits(gs[1]) != bits(hs.center.y) || bits(gs[2]) != bits(hs.center.z) ||
       bits(gs[3]) != bits(hs.radius))
        return fail("bounding sphere differs between host and GPU");
    const AABB hb = compute_aabb(cloud);
    float gb[6];
    ctx->check(vkrt_compute_aabb(ctx->handle(), flat.data(), nv, gb, 0));
    if(gb[0] != hb.min.x || gb[4] != hb.max.y || gb[5] != hb.max.z) return fail("aabb differs between host and GPU");
    const vec3 cen = compute_centroid(std::vector<vec3>{{1.f, 2.f, 3.f}, {3.f, 2.f, 1.f}});
    if(cen.x != 2.f || cen.y != 2.f || cen.z != 2.f) return fail("centroid");

    // PBRPathTracer: an empty scene, two batches, then suspended
    PBRPathTracer::create_info_t info;
    info.seed = 3;
    info.settings.resolution[0] = 32;
    info.settings.resolution[1] = 16;
    info.settings.max_num_batches = 2;
    auto pt = PBRPathTracer::create(ctx, info);
    float zero3[3] = {0.f, 0.f, 0.f};
    uint32_t obj_mat = 0;
    vkrt_material mat{};
    mat.color[0] = mat.color[1] = mat.color[2] = mat.color[3] = 1.f;
    mat.ior = 1.5f;
    mat.roughness = 1.f;
    PBRPathTracer::scene_t scene{zero3, zero3, 0, nullptr, 0, nullptr, &obj_mat, 1, &mat, 1, 1.0f};
    pt->set_scene(scene);
    PBRPathTracer::camera_t cam{{0.f, 0.f, 0.f}, 1.0f, {0.f, 0.f, -1.f}, 0.f, {0.f, 1.f, 0.f}, 1.f};
    for(int f = 0; f < 3; ++f) pt->render_scene(cam);
    if(pt->current_batch() != 2) return fail("batch index");
    if(pt->statistics().size() != 2) return fail("statistics history");
    auto px = pt->pixels();
    if(px.size() != 32 * 16 || !(px[0].radiance[0] > 0.f)) return fail("pixels");
    pt->reset_accumulator();
    if(pt->current_batch() != 0) return fail("reset");

    // the sun of settings_t: direction quat(elevation, azimuth, 0) * (0, 0, 1)
    sunlight_params_t sp;
    sp.spherical_coords[0] = 0.3f;
    sp.spherical_coords[1] = 0.7f;
    const light_t sl = sun_light(sp);
    const double want_dir[3] = {std::cos(0.3) * std::sin(0.7), -std::sin(0.3), std::cos(0.3) * std::cos(0.7)};
    for(int k = 0; k < 3; ++k)
        if(std::fabs(sl.direction[k] - want_dir[k]) > 1e-6) return fail("sun direction");

    // a rough plane (rho 0.5) lit straight down by a delta sun of intensity 2, seen from above: the
    // next-event estimate is f * 2 with f = (0.5 + 0.02) / pi at normal incidence
    float qpos[12] = {-50.f, 0.f, -50.f, 50.f, 0.f, -50.f, 50.f, 0.f, 50.f, -50.f, 0.f, 50.f};
    float qnrm[12] = {0.f, 1.f, 0.f, 0.f, 1.f, 0.f, 0.f, 1.f, 0.f, 0.f, 1.f, 0.f};
    uint32_t qidx[6] = {0, 2, 1, 0, 3, 2}, qobj[2] = {0, 0};
    vkrt_material pm{};
    pm.color[0] = pm.color[1] = pm.color[2] = 0.5f;
    pm.color[3] = 1.f;
    pm.roughness = 1.f;
    pm.ior = 1.5f;
    PBRPathTracer::scene_t plane{qpos, qnrm, 4, qidx, 2, qobj, &obj_mat, 1, &pm, 1, 0.0f};
    PBRPathTracer::create_info_t lit_info;
    lit_info.settings.resolution[0] = lit_info.settings.resolution[1] = 8;
    lit_info.settings.num_samples = 16;
    lit_info.settings.max_trace_depth = 2;
    sunlight_params_t sun;
    sun.color[0] = sun.color[1] = sun.color[2] = 1.f;
    sun.intensity = 2.f;
    sun.spherical_coords[0] = 1.5707963267948966f;
    sun.spherical_coords[1] = 0.f;
    lit_info.settings.sunlight_params = sun;
    auto lit = PBRPathTracer::create(ctx, lit_info);
    lit->set_scene(plane);
    PBRPathTracer::camera_t down{{0.f, 5.f, 0.f}, 0.0698131701f, {0.f, -1.f, 0.f}, 0.f, {0.f, 0.f, -1.f}, 1.f};
    lit->render_scene(down);
    double mean = 0.0;
    for(const auto &p: lit->pixels()) mean += p.radiance[1];
    mean /= 64.0;
    const double want_lit = 0.52 / 3.14159265358979 * 2.0;
    if(std::fabs(mean / want_lit - 1.0) > 0.01) return fail("sun-lit plane");

    // the same plane as an instanced scene: a quarter-size quad placed once with scale 2 (x and z) and a
    // half turn about y
    float hpos[12];
    for(int k = 0; k < 12; ++k) hpos[k] = qpos[k] * 0.5f;
    PBRPathTracer::scene_t half_plane{hpos, qnrm, 4, qidx, 2, qobj, &obj_mat, 1, &pm, 1, 0.0f};
    PBRPathTracer::instance_t place{};
    place.rotation[1] = 1.f;
    place.scale[0] = place.scale[1] = place.scale[2] = 2.f;
    auto placed = PBRPathTracer::create(ctx, lit_info);
    placed->set_scene(half_plane, {0u, 2u}, {place});
    placed->render_scene(down);
    double pmean = 0.0;
    for(const auto &p: placed->pixels()) pmean += p.radiance[1];
    pmean /= 64.0;
    if(std::fabs(pmean / want_lit - 1.0) > 0.01) return fail("sun-lit plane as an instance");

    // the same plane from an OBJ + MTL file through vierkant::model::load_model renders the same
    const auto dir = std::filesystem::temp_directory_path();
    std::ofstream(dir / "vkrt_plane.mtl") << "newmtl grey\nKd 0.5 0.5 0.5\nNs 0\nNi 1.5\n";
    std::ofstream(dir / "vkrt_plane.obj") << "mtllib vkrt_plane.mtl\nv -50 0 -50\nv 50 0 -50\nv 50 0 50\n"
                                             "v -50 0 50\nvn 0 1 0\nusemtl grey\nf 1//1 4//1 3//1 2//1\n";
    std::string why;
    auto file = model::load_model(dir / "vkrt_plane.obj", &why);
    if(!file) return fail(("load_model: " + why).c_str());
    const vkrt_model_view mv = file->view();
    if(mv.n_tri != 2 || mv.n_objects != 1 || mv.n_materials != 1 || file->num_textures() != 0)
        return fail("loaded plane layout");
    if(model::load_model(dir / "vkrt_missing.obj", &why) || why.empty()) return fail("missing file error");
    auto loaded = PBRPathTracer::create(ctx, lit_info);
    loaded->set_scene(file->pt_scene(0.0f));
    loaded->render_scene(down);
    double lmean = 0.0;
    for(const auto &p: loaded->pixels()) lmean += p.radiance[1];
    lmean /= 64.0;
    if(std::fabs(lmean / want_lit - 1.0) > 0.01) return fail("sun-lit plane from an OBJ file");
    std::printf("host api ok (%d hits of %zu rays)\n", n_hit, rays.size());
    return 0;
}
