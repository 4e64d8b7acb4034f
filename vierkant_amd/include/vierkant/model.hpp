// model.hpp - scene files for the path tracer, in the shape of vierkant::model::load_model
// (include/vierkant/model/model_loading.hpp:172): load an OBJ or glTF file, then hand it to
// PBRPathTracer::set_scene. Wraps the C ABI's vkrt_model; host code only.
#pragma once

#include <filesystem>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>

#include "vkrt.h"

namespace vierkant::model
{

//! a loaded scene file: world-space geometry split into objects, materials, textures (RGBA8 with samplers)
class scene_file
{
public:
    //! the arrays (valid while this object lives)
    vkrt_model_view view() const
    {
        vkrt_model_view v{};
        vkrt_model_get(m_model.get(), &v);
        return v;
    }

    //! the path tracer's scene for these arrays with the given sky scale (PBRPathTracer::set_scene)
    vkrt_pt_scene pt_scene(float environment = 1.f) const
    {
        vkrt_pt_scene s{};
        vkrt_model_pt_scene(m_model.get(), environment, &s);
        return s;
    }

    uint32_t num_textures() const { return view().n_textures; }
    uint32_t num_skipped_textures() const { return view().n_skipped_textures; }

private:
    struct deleter
    {
        void operator()(vkrt_model *m) const { vkrt_model_destroy(m); }
    };
    explicit scene_file(vkrt_model *m) : m_model(m) {}
    std::unique_ptr<vkrt_model, deleter> m_model;

    friend std::optional<scene_file> load_model(const std::filesystem::path &path, std::string *error);
};

//! load an OBJ (+ MTL) or glTF 2.0 (.gltf, .glb) file; nullopt on failure, with the reason in *error
inline std::optional<scene_file> load_model(const std::filesystem::path &path, std::string *error = nullptr)
{
    vkrt_model *m = nullptr;
    if(vkrt_model_load(path.string().c_str(), &m) != VKRT_OK)
    {
        if(error) *error = vkrt_model_error();
        return std::nullopt;
    }
    return scene_file(m);
}

}// namespace vierkant::model
