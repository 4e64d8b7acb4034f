/*
 * vkpt_oracle.h - CPU reference integrator (TEST INFRASTRUCTURE ONLY; see vkpt_oracle.c).
 */
#ifndef VKPT_ORACLE_H
#define VKPT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* an RGBA8 UNORM image, rows from the top, 4 bytes per texel, and its sampler: address modes 0 repeat,
 * 1 mirrored repeat, 2 clamp to edge; filters 0 linear, 1 nearest */
typedef struct orc_texture
{
    uint32_t width, height;
    const uint8_t *rgba;
    uint32_t address_mode_u, address_mode_v, min_filter, mag_filter;
} orc_texture;

typedef struct orc_pt_scene
{
    const float *positions; /* 3 per vertex */
    const float *normals;   /* 3 per vertex */
    uint32_t n_vert;
    const uint32_t *indices;
    uint32_t n_tri;
    const uint32_t *tri_object;
    const uint32_t *object_material;
    uint32_t n_objects;
    const float *materials; /* vkrt_material records, 40 words each */
    uint32_t n_materials;
    float environment;
    const float *texcoords; /* 2 per vertex, or NULL (all zero) */
    const orc_texture *textures;
    uint32_t n_textures;
    const float *tangents; /* 3 per vertex, or NULL (all zero) */
    /* two-level scenes (n_instances > 0): the triangles split into meshes at mesh_first (n_meshes + 1
     * ascending offsets), instance k places mesh inst_mesh[k] with the world-to-object rows
     * inst_w2o[12k..] and the transform inst_xf[10k..] = rotation xyzw, translation xyz, scale xyz */
    uint32_t n_meshes;
    const uint32_t *mesh_first;
    uint32_t n_instances;
    const uint32_t *inst_mesh;
    const float *inst_w2o;
    const float *inst_xf;
} orc_pt_scene;

typedef struct orc_pt_params
{
    uint32_t width, height, num_samples, max_depth, seed, disable_material;
    float max_path_beta;
    float eye[3], forward[3], right[3], up[3];
    float tan_half_x, tan_half_y, aperture, focal_distance;
    const float *lights; /* light_t records, 20 words each, already gathered (positive intensity, no Area) */
    uint32_t n_lights, mis_mode, suppress_refractive_caustics;
    float fov_y; /* vertical field of view: the ray cone's spread angle */
    /* test hooks, not reference behaviour: ORC_PT_NO_ROULETTE turns Russian roulette off,
     * ORC_PT_ULP_BIAS moves every transcendental result of the integrator one float up */
    uint32_t test_flags;
} orc_pt_params;

enum
{
    ORC_PT_NO_ROULETTE = 1,
    ORC_PT_ULP_BIAS = 2
};

/* One batch: out8 = 8 floats per pixel (radiance rgb / num_samples, coverage, linear depth, -, object id
 * bits, -), row-major from the top-left. */
void orc_pt_render(const orc_pt_scene *sc, const orc_pt_params *pp, float *out8, int nthreads);
/* vkrt_pt_trace_transmittance on the same scene: rays8 = origin, tmin, direction, tmax; out4 = rgb, 0 */
void orc_pt_transmittance(const orc_pt_scene *sc, const float *rays8, uint32_t n, int cull, float *out4);
void orc_denoise(const uint32_t *px8, uint32_t w, uint32_t h, int denoise, float *out4);
/* test hook: only the first `passes` (1..5) of the five a-trous passes (steps 1, 2, 4, 8, 16) */
void orc_denoise_passes(const uint32_t *px8, uint32_t w, uint32_t h, int passes, float *out4);
/* Test hooks: sample_disney (bsdf_disney.slang:162-296) for n draws on one material record (40 words) at
 * normal N, view V, eta; draw i starts from the RNG state xxhash32(seed, i). out8 per draw: direction xyz,
 * pdf, F rgb, transmission flag. eval_disney (:298-381) for n directions L: out4 = F rgb, pdf. */
void orc_sample_disney_n(const float *mat36, const float *N, const float *V, float eta, uint32_t seed, uint32_t n,
                         float *out8);
void orc_eval_disney_n(const float *mat36, const float *N, const float *V, float eta, const float *L, uint32_t n,
                       float *out4);
/* binary16 rounding (nearest, ties to even) and the exposure/gamma composite of the output image */
float orc_half_value(float x);
void orc_tonemap(const float *in4, size_t n, float exposure, float gamma, float *out4);
/* SampleLevel(u, v, lambda) of one texture through its box-filtered mip chain: rgba in [0, 1] */
void orc_texture_sample(const orc_texture *tex, const float *uv, const float *lambda, size_t n, float *out4);

#ifdef __cplusplus
}
#endif

#endif
