/*
 * vkpt_oracle.c - CPU reference integrator for the path tracer (TEST INFRASTRUCTURE ONLY).
 *
 * A plain-C restatement, pixel by pixel, of the reference's ray-tracing programs for the scenes the
 * tracer supports (absorbing interior media but no scattering or emissive media, no iridescence/dispersion):
 *   raygen trace() + main, ghost-light pass    shaders/slang/raypipeline.slang:58-270 (:201-222)
 *   closest_hit (surface subset), NEE call     shaders/slang/raypipeline.slang:360-680 (:615-619)
 *   homogeneous absorbing medium step          shaders/slang/raypipeline.slang:380-415, :493
 *   miss_default, eval_distant_lights          shaders/slang/raypipeline.slang:728-761
 *   sample_disney, eval_disney + lobes         shaders/slang/ray/bsdf_disney.slang:44-381
 *   analytic lights, MIS, nee_surface          shaders/slang/ray/direct_lighting.slang:32-336
 *   transmittance_test (shadow rays)           shaders/slang/ray/visibility_test.slang:94-165
 *   bsdf / sampling / random / hash   shaders/slang/utils/{bsdf,sampling,random,hash}.slang
 *   environment_white                 shaders/slang/utils/procedural_environment.slang:7-11
 * Closest hits use the oracle's restatement of intersect(Triangle, Ray) (vkoracle.c) with tmin < t < tmax,
 * lowest index on equal t, and back-face culling (det < 0) for rays in air unless the triangle's material
 * is two-sided, through a simple median-split BVH whose box tests are done in double precision on boxes
 * grown by 2^-16 (never culls a triangle the test accepts).
 * Used only by tests/ as the judge of the GPU image (RMSE at equal samples per pixel).
 */
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "vkoracle.h"
#include "vkpt_oracle.h"

typedef struct
{
    float x, y, z;
} v3;

static inline v3 V3(float x, float y, float z) { return (v3){x, y, z}; }
static inline v3 add(v3 a, v3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 scl(v3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
static inline v3 mul(v3 a, v3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 neg(v3 a) { return V3(-a.x, -a.y, -a.z); }
static inline float dot(v3 a, v3 b)
{
    float x = a.x * b.x, y = a.y * b.y, z = a.z * b.z;
    return (x + y) + z;
}
static inline v3 cross(v3 x, v3 y) { return V3(x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y); }
static inline v3 nrmz(v3 v)
{
    float inv = 1.0f / sqrtf(dot(v, v));
    return scl(v, inv);
}
static inline float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
static inline float lerpf(float a, float b, float t) { return a + t * (b - a); }
static inline v3 lerp3(v3 a, v3 b, float t) { return add(a, scl(sub(b, a), t)); }
static inline v3 reflect3(v3 i, v3 n) { return sub(i, scl(n, 2.0f * dot(n, i))); }
static inline float max3(v3 v) { return v.x > v.y ? fmaxf(v.x, v.z) : fmaxf(v.y, v.z); }
static inline v3 refract3(v3 i, v3 n, float eta)
{
    float ndi = dot(n, i);
    float k = 1.0f - eta * eta * (1.0f - ndi * ndi);
    if(k < 0.0f) return V3(0, 0, 0);
    return sub(scl(i, eta), scl(n, eta * ndi + sqrtf(k)));
}

static const float PI = 3.1415926535897932384626433832795f;
static const float ONE_OVER_PI = 0.31830988618379067153776752674503f;
static const float EPS = 1e-4f, PDF_EPS = 1e-3f, DEFAULT_TMAX = 10000.0f;

/* ---- RNG (hash.slang, random.slang, sampling.slang) ------------------------------------------ */
static inline float rnd(uint32_t *s)
{
    *s = 1664525u * *s + 1013904223u;
    return (float)(*s & 0x00FFFFFFu) / (float)0x01000000;
}

static inline uint32_t brev(uint32_t x)
{
    uint32_t r = 0;
    for(int i = 0; i < 32; ++i) r |= ((x >> i) & 1u) << (31 - i);
    return r;
}

/* ---- sampling -------------------------------------------------------------------------------- */
typedef struct
{
    v3 tx, ty, n;
} frame;

static frame local_frame(v3 d)
{
    float len2 = d.x * d.x + d.y * d.y;
    float len = sqrtf(len2);
    frame f;
    f.tx = len2 > 0.f ? V3(-d.y / len, d.x / len, 0.f / len) : V3(1, 0, 0);
    f.ty = cross(d, f.tx);
    f.n = d;
    return f;
}
static inline v3 to_local(frame f, v3 v) { return V3(dot(f.tx, v), dot(f.ty, v), dot(f.n, v)); }
static inline v3 to_world(frame f, v3 v) { return add(add(scl(f.tx, v.x), scl(f.ty, v.y)), scl(f.n, v.z)); }
/* Sensitivity hook (test infrastructure, not reference behaviour): with g_ulp_bias set, every
 * transcendental of the integrator (cos, sin, tan, atan, asin, pow, log2, exp) returns the next float above
 * the library's result, so a test can measure how far a one-ulp difference between two math libraries moves
 * an image. Set only for the duration of one orc_pt_render call (not for concurrent renders). */
static int g_ulp_bias = 0;
static inline float ulp_up(float r) { return g_ulp_bias ? nextafterf(r, INFINITY) : r; }
static inline float tcos(float x) { return ulp_up(cosf(x)); }
static inline float tsin(float x) { return ulp_up(sinf(x)); }
static inline float ttan(float x) { return ulp_up(tanf(x)); }
static inline float tatan(float x) { return ulp_up(atanf(x)); }
static inline float tasin(float x) { return ulp_up(asinf(x)); }
static inline float tpow(float x, float y) { return ulp_up(powf(x, y)); }
static inline float tlog2(float x) { return ulp_up(log2f(x)); }
static inline float texp(float x) { return ulp_up(expf(x)); }

static inline v3 sph(float s, float c, float phi) { return V3(s * tcos(phi), s * tsin(phi), c); }

static v3 hemi_cos(float xi_x, float xi_y)
{
    float phi = 2.0f * PI * xi_y;
    float c = sqrtf(fmaxf(1.0f - xi_x, 0.0f));
    float s = sqrtf(fmaxf(1.0f - c * c, 0.0f));
    return sph(s, c, phi);
}

static v3 ggx_vndf(float xi_x, float xi_y, v3 wi, float a)
{
    v3 w = nrmz(V3(wi.x * a, wi.y * a, wi.z));
    float phi = (2.0f * xi_x - 1.0f) * PI;
    float z = fmaf(1.0f - xi_y, 1.0f + w.z, -w.z);
    float st = sqrtf(clampf(1.0f - z * z, 0.0f, 1.0f));
    v3 c = V3(st * tcos(phi), st * tsin(phi), z);
    v3 m = add(c, w);
    return nrmz(V3(m.x * a, m.y * a, m.z));
}

static v3 gtr1_sample(float xi_x, float xi_y, float r)
{
    float a = fmaxf(0.001f, r), a2 = a * a;
    float phi = xi_y * 2.0f * PI;
    float c = sqrtf((1.0f - tpow(a2, 1.0f - xi_x)) / (1.0f - a2));
    float s = clampf(sqrtf(1.0f - c * c), 0.0f, 1.0f);
    return sph(s, c, phi);
}

/* ---- bsdf.slang ------------------------------------------------------------------------------ */
static inline float schlick(float u)
{
    float m = clampf(1.0f - u, 0.0f, 1.0f);
    return m * m * m * m * m;
}
static float diel_fresnel(float ci, float eta)
{
    float s2 = eta * eta * (1.0f - ci * ci);
    if(s2 >= 1.0f) return 1.0f;
    float ct = sqrtf(fmaxf(1.0f - s2, 0.0f));
    float rs = (eta * ct - ci) / (eta * ct + ci);
    float rp = (eta * ci - ct) / (eta * ci + ct);
    return 0.5f * (rs * rs + rp * rp);
}
static inline float gtr2(float nh, float a)
{
    float a2 = a * a;
    float t = 1.0f + (a2 - 1.0f) * nh * nh;
    return a2 / (PI * t * t);
}
static inline float smith(float nv, float ag)
{
    float a = ag * ag, b = nv * nv;
    return 1.0f / (nv + sqrtf(a + b - a * b));
}

typedef struct
{
    v3 color;
    v3 emission;
    float emission_strength;
    float metal, rough, trans, ior, cc, cc_rough, sheen_rough, dt;
    v3 sheen, dt_color;
    int two_sided;
    float alpha;        /* color.a */
    uint32_t tex;       /* base color texture + 1, 0 = none */
    uint32_t tex_normal, tex_emission, tex_arm, tex_transmission, tex_dt, tex_dtc; /* further slots, same code */
    int alpha_mode;     /* 0 opaque, 1 MASK (cutoff), 2 BLEND (uniform draw): the any-hit test */
    float alpha_cutoff;
    int vol;            /* an interior medium (attenuation_distance > 0) */
    v3 sigma;           /* its extinction -log(attenuation_color) / attenuation_distance (raypipeline.slang:493) */
} mat_t;

typedef struct
{
    v3 dir;
    float pdf;
    v3 F;
    int transmission;
} bs_t;

static v3 ev_diel_refl(float r, v3 color, float eta, v3 V, v3 N, v3 L, v3 H, float *pdf)
{
    *pdf = 0.0f;
    if(dot(N, L) <= 0.0f) return V3(0, 0, 0);
    float F = diel_fresnel(dot(V, H), eta);
    float D = gtr2(dot(N, H), r);
    *pdf = 0.5f * D * smith(fabsf(dot(N, V)), r) * F;
    float G = smith(fabsf(dot(N, L)), r) * smith(fabsf(dot(N, V)), r);
    return scl(scl(scl(color, F), D), G);
}

static v3 ev_diel_refr(float r, v3 color, float eta, v3 V, v3 N, v3 L, v3 H, float *pdf)
{
    *pdf = 0.0f;
    if(dot(N, L) >= 0.0f) return V3(0, 0, 0);
    float ds = dot(L, H) + dot(V, H) * eta;
    if(fabsf(ds) < 1e-7f) return V3(0, 0, 0);
    float F = diel_fresnel(fabsf(dot(V, H)), eta);
    float D = gtr2(dot(N, H), r);
    *pdf = 2.0f * smith(fabsf(dot(N, V)), r) * D * fabsf(dot(V, H)) * (1.0f - F) * fabsf(dot(L, H)) * eta * eta /
           (ds * ds);
    float G = smith(fabsf(dot(N, L)), r) * smith(fabsf(dot(N, V)), r);
    v3 c = scl(color, 1.0f - F);
    c = scl(c, D);
    c = scl(c, G);
    c = scl(c, fabsf(dot(V, H)));
    c = scl(c, fabsf(dot(L, H)));
    c = scl(c, 4.0f);
    c = scl(c, eta);
    c = scl(c, eta);
    float dd = ds * ds;
    return V3(c.x / dd, c.y / dd, c.z / dd);
}

static v3 ev_spec(float r, v3 cs, v3 V, v3 N, v3 L, v3 H, float *pdf)
{
    *pdf = 0.0f;
    if(dot(N, L) <= 0.0f) return V3(0, 0, 0);
    float D = gtr2(dot(N, H), r);
    *pdf = 0.5f * D * smith(fabsf(dot(N, V)), r);
    float FH = schlick(dot(L, H));
    v3 F = lerp3(cs, V3(1, 1, 1), FH);
    float G = smith(fabsf(dot(N, L)), r) * smith(fabsf(dot(N, V)), r);
    return scl(scl(F, D), G);
}

static v3 ev_cc(float ccr, float cc, v3 V, v3 N, v3 L, v3 H, float *pdf)
{
    *pdf = 0.0f;
    if(dot(N, L) <= 0.0f) return V3(0, 0, 0);
    float D = gtr2(dot(N, H), ccr);
    *pdf = D * dot(N, H) / (4.0f * dot(V, H));
    float FH = schlick(dot(L, H));
    float F = lerpf(0.04f, 1.0f, FH);
    float G = smith(dot(N, L), 0.25f) * smith(fabsf(dot(N, V)), 0.25f);
    float x = 0.25f * cc * F * D * G;
    return V3(x, x, x);
}

static v3 ev_diffuse(float r, float sr, v3 color, v3 sheen, float metal, v3 V, v3 N, v3 L, v3 H, float *pdf)
{
    *pdf = 0.0f;
    float NoL = dot(N, L);
    if(NoL <= 0.0f) return V3(0, 0, 0);
    *pdf = NoL * ONE_OVER_PI;
    float LoH = dot(L, H);
    float FL = schlick(NoL), FV = schlick(dot(N, V)), FH = schlick(LoH);
    float Fd90 = 0.5f + 2.0f * LoH * LoH * r;
    float Fd = lerpf(1.0f, Fd90, FL) * lerpf(1.0f, Fd90, FV);
    v3 Fs = scl(sheen, FH * sr);
    return scl(scl(add(scl(color, Fd), Fs), ONE_OVER_PI), 1.0f - metal);
}

static bs_t sample_disney(mat_t m, v3 N, v3 V, float eta, uint32_t *rng)
{
    bs_t ret;
    ret.pdf = 0.0f;
    ret.F = V3(0, 0, 0);
    ret.transmission = 0;
    ret.dir = V3(0, 0, 0);
    float diffuse_ratio = 0.5f * (1.0f - m.metal);
    float trans_weight = (1.0f - m.metal) * m.trans;
    m.rough = fmaxf(0.001f, m.rough * m.rough);
    m.sheen_rough = fmaxf(0.001f, m.sheen_rough);
    v3 spec_color = lerp3(V3(0.08f, 0.08f, 0.08f), m.color, m.metal);
    frame fr = local_frame(N);
    float xi_x = rnd(rng), xi_y = rnd(rng);
    v3 H = to_world(fr, ggx_vndf(xi_x, xi_y, to_local(fr, V), m.rough));
    if(dot(V, H) < 0.0f) H = neg(H);
    if(rnd(rng) < trans_weight)
    {
        v3 R = reflect3(neg(V), H);
        float F = diel_fresnel(fabsf(dot(R, H)), eta);
        if(xi_y < F)
        {
            ret.dir = nrmz(R);
            (void)rnd(rng);
            ret.F = ev_diel_refl(m.rough, m.color, eta, V, N, ret.dir, H, &ret.pdf);
            ret.pdf *= 1.0f - 0.0f;
        }
        else
        {
            v3 rf = refract3(neg(V), H, eta);
            if(dot(rf, rf) < 1e-7f)
            {
                ret.dir = nrmz(reflect3(neg(V), H));
                ret.F = ev_diel_refl(m.rough, m.color, eta, V, N, ret.dir, H, &ret.pdf);
            }
            else
            {
                ret.dir = nrmz(rf);
                ret.F = ev_diel_refr(m.rough, m.color, eta, V, N, ret.dir, H, &ret.pdf);
                ret.transmission = 1;
            }
        }
        ret.F = scl(ret.F, trans_weight);
        ret.pdf *= trans_weight;
    }
    else
    {
        if(rnd(rng) < diffuse_ratio)
        {
            float dt = m.dt;
            if(dt > 0.0f && rnd(rng) < dt)
            {
                v3 l = hemi_cos(xi_x, xi_y);
                ret.dir = to_world(fr, V3(l.x, l.y, -l.z));
                ret.F = scl(m.dt_color, dt * ONE_OVER_PI);
                ret.pdf = l.z * ONE_OVER_PI * diffuse_ratio * dt;
                ret.transmission = 1;
            }
            else
            {
                ret.dir = to_world(fr, hemi_cos(xi_x, xi_y));
                v3 Hd = nrmz(add(ret.dir, V));
                ret.F = scl(ev_diffuse(m.rough, m.sheen_rough, m.color, m.sheen, m.metal, V, N, ret.dir, Hd, &ret.pdf),
                            1.0f - dt);
                ret.pdf *= diffuse_ratio * (1.0f - dt);
            }
        }
        else
        {
            float psr = 1.0f / (1.0f + m.cc);
            if(rnd(rng) < psr)
            {
                ret.dir = nrmz(reflect3(neg(V), H));
                (void)rnd(rng);
                ret.F = ev_spec(m.rough, spec_color, V, N, ret.dir, H, &ret.pdf);
                ret.pdf *= psr * (1.0f - diffuse_ratio) * (1.0f - 0.0f);
            }
            else
            {
                v3 Hc = to_world(fr, gtr1_sample(xi_x, xi_y, lerpf(0.1f, 0.001f, m.cc)));
                if(dot(V, Hc) < 0.0f) Hc = neg(Hc);
                ret.dir = nrmz(reflect3(neg(V), Hc));
                ret.F = ev_cc(m.cc_rough, m.cc, V, N, ret.dir, Hc, &ret.pdf);
                ret.pdf *= (1.0f - psr) * (1.0f - diffuse_ratio);
            }
        }
        ret.F = scl(ret.F, 1.0f - trans_weight);
        ret.pdf *= (1.0f - trans_weight);
    }
    if(ret.pdf != ret.pdf) ret.pdf = 0.0f;
    return ret;
}

/* a vkrt_material record (40 words) as the integrator's material */
static mat_t unpack_material(const float *x, uint32_t n_textures)
{
    mat_t mt;
    memset(&mt, 0, sizeof(mt));
    mt.color = V3(x[0], x[1], x[2]);
    mt.emission = V3(x[4], x[5], x[6]);
    mt.emission_strength = x[7];
    mt.metal = x[8];
    mt.rough = x[9];
    mt.trans = x[10];
    mt.ior = x[11];
    mt.cc = x[12];
    mt.cc_rough = x[13];
    mt.sheen_rough = x[14];
    mt.dt = x[15];
    mt.sheen = V3(x[16], x[17], x[18]);
    mt.dt_color = V3(x[20], x[21], x[22]);
    uint32_t ts, tex, mode;
    memcpy(&ts, &x[24], 4);
    mt.two_sided = ts != 0;
    mt.alpha = x[3];
    memcpy(&tex, &x[25], 4);
    memcpy(&mode, &x[26], 4);
    mt.tex = tex <= n_textures ? tex : 0;
    mt.alpha_mode = mode <= 2 ? (int)mode : 0;
    mt.alpha_cutoff = x[27];
    uint32_t slot[6];
    memcpy(slot, &x[28], sizeof(slot));
    for(int k = 0; k < 6; ++k) slot[k] = slot[k] <= n_textures ? slot[k] : 0;
    mt.tex_normal = slot[0];
    mt.tex_emission = slot[1];
    mt.tex_arm = slot[2];
    mt.tex_transmission = slot[3];
    mt.tex_dt = slot[4];
    mt.tex_dtc = slot[5];
    if(x[37] > 0.0f)
    {
        mt.vol = 1;
        mt.sigma = V3(-logf(x[34]) / x[37], -logf(x[35]) / x[37], -logf(x[36]) / x[37]);
    }
    return mt;
}

static v3 env_white(v3 d)
{
    v3 sky = V3(1.3f, 1.3f, 1.3f), hor = V3(0.6f, 0.6f, 0.6f), gr = V3(0.1f, 0.1f, 0.1f);
    return d.y > 0.f ? lerp3(hor, sky, d.y) : lerp3(hor, gr, -d.y);
}

static float round_half(float x)
{
    uint32_t b;
    memcpy(&b, &x, 4);
    uint32_t sign = b & 0x80000000u, a = b & 0x7fffffffu;
    if(a >= 0x47800000u || a < 0x38800000u) return x;
    uint32_t lsb = (a >> 13) & 1u;
    uint32_t r = ((a + 0xfffu + lsb) & ~0x1fffu) | sign;
    float out;
    memcpy(&out, &r, 4);
    return out;
}

/* ---- textures ---------------------------------------------------------------------------------
 * RGBA8 UNORM images with a full mip chain, sampled as the path tracer's Sampler2D (raypipeline.slang:
 * 312-321 SampleLevel; sampler defaults of Material.hpp:175-179: linear filtering, linear mip blending,
 * repeat addressing). Each level halves the previous one (at least one texel) with a 2 x 2 box filter whose
 * coordinates clamp at an odd edge, rounded to nearest: (a + b + c + d + 2) / 4. */
#define ORC_MAX_LEVELS 16
typedef struct
{
    uint32_t w, h, levels;
    uint32_t mode_u, mode_v, min_nearest, mag_nearest;
    uint8_t *lv[ORC_MAX_LEVELS];
} otex_t;

static void otex_build(otex_t *t, const orc_texture *src)
{
    memset(t, 0, sizeof(*t));
    t->w = src->width;
    t->h = src->height;
    t->mode_u = src->address_mode_u;
    t->mode_v = src->address_mode_v;
    t->min_nearest = src->min_filter == 1;
    t->mag_nearest = src->mag_filter == 1;
    size_t bytes = (size_t)t->w * t->h * 4;
    t->lv[0] = (uint8_t *)malloc(bytes ? bytes : 4);
    memcpy(t->lv[0], src->rgba, bytes);
    t->levels = 1;
    uint32_t w = t->w, h = t->h;
    while((w > 1 || h > 1) && t->levels < ORC_MAX_LEVELS)
    {
        const uint32_t nw = w > 1 ? w / 2 : 1, nh = h > 1 ? h / 2 : 1;
        const uint8_t *p = t->lv[t->levels - 1];
        uint8_t *q = (uint8_t *)malloc((size_t)nw * nh * 4);
        for(uint32_t y = 0; y < nh; ++y)
            for(uint32_t x = 0; x < nw; ++x)
            {
                const uint32_t xs[2] = {2 * x < w - 1 ? 2 * x : w - 1, 2 * x + 1 < w - 1 ? 2 * x + 1 : w - 1};
                const uint32_t ys[2] = {2 * y < h - 1 ? 2 * y : h - 1, 2 * y + 1 < h - 1 ? 2 * y + 1 : h - 1};
                for(int c = 0; c < 4; ++c)
                {
                    uint32_t sum = 2;
                    for(int a = 0; a < 2; ++a)
                        for(int b = 0; b < 2; ++b) sum += p[4 * ((size_t)ys[a] * w + xs[b]) + c];
                    q[4 * ((size_t)y * nw + x) + c] = (uint8_t)(sum / 4);
                }
            }
        t->lv[t->levels++] = q;
        w = nw;
        h = nh;
    }
}

static void otex_free(otex_t *t)
{
    for(uint32_t l = 0; l < t->levels; ++l) free(t->lv[l]);
}

/* Vulkan's address modes for texel index i of a level n wide: repeat, mirrored repeat, clamp to edge */
static int otex_address(long long i, int n, uint32_t mode)
{
    if(mode == 2) return (int)(i < 0 ? 0 : (i >= n ? n - 1 : i));
    if(mode == 1)
    {
        long long m = 2ll * n, r = ((i % m) + m) % m;
        return (int)(r < n ? r : m - 1 - r);
    }
    return (int)(((i % n) + n) % n);
}

/* one level: bilinear (texel centres at (i + 0.5) / size) or nearest (texel floor(u size)) */
static void otex_level(const otex_t *t, uint32_t l, float u, float v, int nearest, float out[4])
{
    const int w = (int)(t->w >> l ? t->w >> l : 1), h = (int)(t->h >> l ? t->h >> l : 1);
    const uint8_t *p = t->lv[l];
    if(nearest)
    {
        const int x = otex_address((long long)floorf(u * (float)w), w, t->mode_u);
        const int y = otex_address((long long)floorf(v * (float)h), h, t->mode_v);
        for(int c = 0; c < 4; ++c) out[c] = (float)p[4 * ((size_t)y * w + x) + c] / 255.0f;
        return;
    }
    const float fx = u * (float)w - 0.5f, fy = v * (float)h - 0.5f;
    const float x0f = floorf(fx), y0f = floorf(fy);
    const float ax = fx - x0f, ay = fy - y0f;
    const long long xl = (long long)x0f, yl = (long long)y0f;
    const int x0 = otex_address(xl, w, t->mode_u), x1 = otex_address(xl + 1, w, t->mode_u);
    const int y0 = otex_address(yl, h, t->mode_v), y1 = otex_address(yl + 1, h, t->mode_v);
    for(int c = 0; c < 4; ++c)
    {
        const float c00 = (float)p[4 * ((size_t)y0 * w + x0) + c] / 255.0f;
        const float c10 = (float)p[4 * ((size_t)y0 * w + x1) + c] / 255.0f;
        const float c01 = (float)p[4 * ((size_t)y1 * w + x0) + c] / 255.0f;
        const float c11 = (float)p[4 * ((size_t)y1 * w + x1) + c] / 255.0f;
        out[c] = lerpf(lerpf(c00, c10, ax), lerpf(c01, c11, ax), ay);
    }
}

/* SampleLevel: lambda clamped to [0, levels - 1] (NaN and -inf give level 0), linear between the two
 * levels around it, the min filter where lambda > 0 and the mag filter elsewhere; coordinates beyond 1e9 in
 * magnitude carry no fraction and read as (0, 0) */
static void otex_sample(const otex_t *t, float u, float v, float lambda, float out[4])
{
    const float top = (float)(t->levels - 1);
    const int nearest = lambda > 0.0f ? (int)t->min_nearest : (int)t->mag_nearest;
    if(!(lambda > 0.0f)) lambda = 0.0f;
    if(lambda > top) lambda = top;
    if(!(fabsf(u) < 1e9f) || !(fabsf(v) < 1e9f)) u = v = 0.0f;
    const uint32_t l0 = (uint32_t)lambda;
    const float f = lambda - (float)l0;
    otex_level(t, l0, u, v, nearest, out);
    if(f == 0.0f) return;
    float b[4];
    otex_level(t, l0 + 1 < t->levels ? l0 + 1 : t->levels - 1, u, v, nearest, b);
    for(int c = 0; c < 4; ++c) out[c] = lerpf(out[c], b[c], f);
}

void orc_texture_sample(const orc_texture *tex, const float *uv, const float *lambda, size_t n, float *out4)
{
    otex_t t;
    otex_build(&t, tex);
    for(size_t i = 0; i < n; ++i) otex_sample(&t, uv[2 * i], uv[2 * i + 1], lambda[i], out4 + 4 * i);
    otex_free(&t);
}

/* what the any-hit stages need of the scene: the alpha-cutout test of MASK materials (raypipeline.slang:
 * 686-720 for closest-hit rays, visibility_test.slang:136-155 for shadow rays) */
typedef struct
{
    const mat_t *mats;
    const uint32_t *mat_of_prim;
    const uint32_t *indices;
    const float *uv; /* 2 per vertex */
    const otex_t *tex;
} alpha_ctx;

/* 1 when the candidate is kept: color.a (times the base color texture's alpha at level 0, at the hit's
 * interpolated coordinates) is not below the cutoff (MASK) or a uniform draw (BLEND). disable_material stands
 * the color in by (0.8, 1) for closest-hit rays, which draw once per trace from a copy of the path's state
 * (blend_xi); the shadow-ray test reads the material as it is and draws per candidate, here rnd() of a copy
 * of xxhash32(blend_seed, triangle), so that no traversal order enters (per_candidate). */
static int alpha_keep(const alpha_ctx *ac, uint32_t prim, float bu, float bv, int disable_material, float blend_xi,
                      uint32_t blend_seed, int per_candidate)
{
    const mat_t *m = &ac->mats[ac->mat_of_prim[prim]];
    if(!m->alpha_mode) return 1;
    float a = m->alpha;
    if(disable_material) a = 1.0f;
    else if(m->tex)
    {
        const uint32_t *ix = ac->indices + 3 * (size_t)prim;
        const float *t0 = ac->uv + 2 * (size_t)ix[0], *t1 = ac->uv + 2 * (size_t)ix[1], *t2 = ac->uv + 2 * (size_t)ix[2];
        const float b0 = 1.0f - bu - bv;
        const float u = (t0[0] * b0 + t1[0] * bu) + t2[0] * bv, v = (t0[1] * b0 + t1[1] * bu) + t2[1] * bv;
        float c[4];
        otex_sample(&ac->tex[m->tex - 1], u, v, 0.0f, c);
        a = a * c[3];
    }
    if(m->alpha_mode == 1) return !(a < m->alpha_cutoff);
    if(per_candidate)
    {
        uint32_t st = orc_xxhash32(blend_seed, prim);
        blend_xi = rnd(&st);
    }
    return !(a < blend_xi);
}

/* ---- a CPU BVH with exact closest-hit semantics ----------------------------------------------- */
typedef struct
{
    double lo[3], hi[3];
    int32_t left, right; /* children (node index), or -1 for a leaf */
    uint32_t first, count;
} cnode;

typedef struct
{
    const float *tris; /* 9 floats per triangle */
    const uint8_t *two_sided;
    uint32_t n;
    cnode *nodes;
    uint32_t n_nodes;
    uint32_t *order;
    const alpha_ctx *ac; /* the any-hit test of masked materials */
} cbvh;

static void tri_box(const float *t, double lo[3], double hi[3])
{
    for(int a = 0; a < 3; ++a)
    {
        double mn = fmin(fmin(t[a], t[3 + a]), t[6 + a]), mx = fmax(fmax(t[a], t[3 + a]), t[6 + a]);
        double pad = (fabs(mn) > fabs(mx) ? fabs(mn) : fabs(mx)) * (1.0 / 65536.0) + 1e-30;
        lo[a] = mn - pad;
        hi[a] = mx + pad;
    }
}

static uint32_t build_rec(cbvh *b, uint32_t first, uint32_t count, const double *cent)
{
    uint32_t id = b->n_nodes++;
    cnode *nd = &b->nodes[id];
    for(int a = 0; a < 3; ++a)
    {
        nd->lo[a] = INFINITY;
        nd->hi[a] = -INFINITY;
    }
    for(uint32_t k = first; k < first + count; ++k)
    {
        double lo[3], hi[3];
        tri_box(b->tris + 9 * (size_t)b->order[k], lo, hi);
        for(int a = 0; a < 3; ++a)
        {
            nd->lo[a] = fmin(nd->lo[a], lo[a]);
            nd->hi[a] = fmax(nd->hi[a], hi[a]);
        }
    }
    if(count <= 4)
    {
        nd->left = nd->right = -1;
        nd->first = first;
        nd->count = count;
        return id;
    }
    double clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for(uint32_t k = first; k < first + count; ++k)
        for(int a = 0; a < 3; ++a)
        {
            double c = cent[3 * (size_t)b->order[k] + a];
            clo[a] = fmin(clo[a], c);
            chi[a] = fmax(chi[a], c);
        }
    int axis = 0;
    for(int a = 1; a < 3; ++a)
        if(chi[a] - clo[a] > chi[axis] - clo[axis]) axis = a;
    double split = 0.5 * (clo[axis] + chi[axis]);
    uint32_t i = first, j = first + count;
    while(i < j)
    {
        if(cent[3 * (size_t)b->order[i] + axis] < split) ++i;
        else
        {
            --j;
            uint32_t t = b->order[i];
            b->order[i] = b->order[j];
            b->order[j] = t;
        }
    }
    uint32_t mid = i;
    if(mid == first || mid == first + count) mid = first + count / 2;
    int32_t l = (int32_t)build_rec(b, first, mid - first, cent);
    int32_t r = (int32_t)build_rec(b, mid, first + count - mid, cent);
    nd = &b->nodes[id];
    nd->left = l;
    nd->right = r;
    nd->first = nd->count = 0;
    return id;
}

static int box_hit(const cnode *nd, const double o[3], const double inv[3], double tmin, double tmax)
{
    double t0 = -INFINITY, t1 = INFINITY;
    for(int a = 0; a < 3; ++a)
    {
        double ta = (nd->lo[a] - o[a]) * inv[a], tb = (nd->hi[a] - o[a]) * inv[a];
        if(ta != ta) ta = -INFINITY; /* 0 * inf on a slab plane: inside */
        if(tb != tb) tb = INFINITY;
        double lo = fmin(ta, tb), hi = fmax(ta, tb);
        t0 = fmax(t0, lo);
        t1 = fmin(t1, hi);
    }
    t0 -= fabs(t0) * 1e-6;
    t1 += fabs(t1) * 1e-6;
    return t0 <= t1 && t1 >= tmin && t0 <= tmax;
}

/* the closest record so far: t (tmax before any hit), barycentrics, triangle and instance */
typedef struct
{
    float t, u, v;
    uint32_t prim, inst;
} chit_t;

/* closest hit with tmin < t < h->t over one BVH whose triangles belong to instance `inst` (0 for a flat
 * scene): the smallest t wins, then the lowest instance, then the lowest triangle; back faces culled when
 * `cull`; masked candidates whose alpha falls below their cutoff are ignored (any-hit IgnoreHit) */
static void closest(const cbvh *b, v3 o, v3 d, float tmin, uint32_t inst, int cull, int disable_material,
                    float blend_xi, chit_t *h)
{
    double od[3] = {o.x, o.y, o.z}, inv[3] = {1.0 / d.x, 1.0 / d.y, 1.0 / d.z};
    uint32_t stack[128];
    int sp = 0;
    if(b->n_nodes == 0) return;
    stack[sp++] = 0;
    while(sp)
    {
        const cnode *nd = &b->nodes[stack[--sp]];
        if(!box_hit(nd, od, inv, tmin, h->t)) continue;
        if(nd->left < 0)
        {
            for(uint32_t k = nd->first; k < nd->first + nd->count; ++k)
            {
                uint32_t prim = b->order[k];
                const float *tr = b->tris + 9 * (size_t)prim;
                float t, u, v;
                if(!orc_intersect_tri(tr, &o.x, &d.x, &t, &u, &v)) continue;
                if(cull && !b->two_sided[prim])
                {
                    v3 e1 = V3(tr[3] - tr[0], tr[4] - tr[1], tr[5] - tr[2]);
                    v3 e2 = V3(tr[6] - tr[0], tr[7] - tr[1], tr[8] - tr[2]);
                    if(dot(e1, cross(d, e2)) < 0.f) continue;
                }
                if(!(t > tmin)) continue;
                if(t < h->t ||
                   (t == h->t && h->prim != 0xFFFFFFFFu && (inst < h->inst || (inst == h->inst && prim < h->prim))))
                {
                    if(!alpha_keep(b->ac, prim, u, v, disable_material, blend_xi, 0u, 0)) continue;
                    h->t = t;
                    h->u = u;
                    h->v = v;
                    h->prim = prim;
                    h->inst = inst;
                }
            }
        }
        else
        {
            stack[sp++] = (uint32_t)nd->left;
            stack[sp++] = (uint32_t)nd->right;
        }
    }
}

/* transmittance_test (visibility_test.slang:94-165): 0 when an opaque triangle (not cut out by its alpha
 * mask) lies in (tmin, tmax), else the product of transmission * color over the transmissive hits in
 * ascending (t, instance, triangle) order; mat_of_prim maps a triangle to its material. With `cull`, back
 * faces (det < 0 in the triangle's own space) of one-sided triangles are skipped. transmit_collect gathers
 * one BVH's hits (instance `inst`) and returns 1 when an opaque hit blocks the ray; transmit_product sorts
 * and multiplies. */
#define ORC_MAX_TRANSMISSIVE 4096
typedef struct
{
    float t;
    uint32_t prim, inst;
    int front; /* the ray enters the triangle's volume here (det > 0 in the triangle's own space) */
} thit_t;

static int transmit_collect(const cbvh *b, const mat_t *mats, const uint32_t *mat_of_prim, v3 o, v3 d, float tmin,
                            float tmax, uint32_t blend_seed, int per_candidate, int cull, uint32_t inst, thit_t *hits,
                            uint32_t *n_hits)
{
    double od[3] = {o.x, o.y, o.z}, inv[3] = {1.0 / d.x, 1.0 / d.y, 1.0 / d.z};
    uint32_t nh = *n_hits, stack[128];
    int sp = 0;
    if(b->n_nodes) stack[sp++] = 0;
    while(sp)
    {
        const cnode *nd = &b->nodes[stack[--sp]];
        if(!box_hit(nd, od, inv, tmin, tmax)) continue;
        if(nd->left >= 0)
        {
            stack[sp++] = (uint32_t)nd->left;
            stack[sp++] = (uint32_t)nd->right;
            continue;
        }
        for(uint32_t k = nd->first; k < nd->first + nd->count; ++k)
        {
            uint32_t prim = b->order[k];
            const float *tr = b->tris + 9 * (size_t)prim;
            float t, u, v;
            if(!orc_intersect_tri(tr, &o.x, &d.x, &t, &u, &v)) continue;
            if(cull && !b->two_sided[prim])
            {
                v3 e1 = V3(tr[3] - tr[0], tr[4] - tr[1], tr[5] - tr[2]);
                v3 e2 = V3(tr[6] - tr[0], tr[7] - tr[1], tr[8] - tr[2]);
                if(dot(e1, cross(d, e2)) < 0.f) continue;
            }
            if(!(t > tmin && t < tmax)) continue;
            if(!(mats[mat_of_prim[prim]].trans > 0.0f))
            {
                /* alpha-cutout surfaces let the shadow ray through (visibility_test.slang:145-150) */
                if(!alpha_keep(b->ac, prim, u, v, 0, 0.0f, blend_seed, per_candidate)) continue;
                *n_hits = nh;
                return 1;
            }
            if(nh == ORC_MAX_TRANSMISSIVE) abort();
            {
                v3 e1 = V3(tr[3] - tr[0], tr[4] - tr[1], tr[5] - tr[2]);
                v3 e2 = V3(tr[6] - tr[0], tr[7] - tr[1], tr[8] - tr[2]);
                hits[nh].front = dot(e1, cross(d, e2)) > 0.f;
            }
            hits[nh].t = t;
            hits[nh].prim = prim;
            hits[nh].inst = inst;
            ++nh;
        }
    }
    *n_hits = nh;
    return 0;
}

static int thit_before(const thit_t *a, const thit_t *b)
{
    return a->t < b->t || (a->t == b->t && (a->inst < b->inst || (a->inst == b->inst && a->prim < b->prim)));
}

static v3 transmit_product(const mat_t *mats, const uint32_t *mat_of_prim, thit_t *hits, uint32_t nh)
{
    for(uint32_t a = 1; a < nh; ++a)
    {
        thit_t x = hits[a];
        uint32_t c = a;
        while(c > 0 && thit_before(&x, &hits[c - 1]))
        {
            hits[c] = hits[c - 1];
            --c;
        }
        hits[c] = x;
    }
    /* Beer-Lambert optical depth over the transmissive hits with an interior medium, a signed sum (entry -,
     * exit +; visibility_test.slang:123-129) taken in the same order, clamped at 0 */
    v3 T = V3(1, 1, 1), D = V3(0, 0, 0);
    for(uint32_t a = 0; a < nh; ++a)
    {
        const mat_t *m = &mats[mat_of_prim[hits[a].prim]];
        T = V3(T.x * (m->trans * m->color.x), T.y * (m->trans * m->color.y), T.z * (m->trans * m->color.z));
        if(m->vol)
        {
            const float t = hits[a].t;
            const int f = hits[a].front;
            D = V3(D.x + (f ? -m->sigma.x * t : m->sigma.x * t), D.y + (f ? -m->sigma.y * t : m->sigma.y * t),
                   D.z + (f ? -m->sigma.z * t : m->sigma.z * t));
        }
    }
    return V3(T.x * expf(-fmaxf(D.x, 0.f)), T.y * expf(-fmaxf(D.y, 0.f)), T.z * expf(-fmaxf(D.z, 0.f)));
}

/* ---- two-level scenes (the instance level of RayBuilder's top-level structure, src/RayBuilder.cpp:386-781):
 * one BVH per mesh over its triangles in object space; instance k places mesh[k] with the world-to-object
 * rows w2o[12k..] (entry_t::inv_transform) and the transform xf[10k..] = rotation xyzw, translation xyz,
 * scale xyz (entry_t::transform) */
typedef struct
{
    uint32_t n_inst;
    const uint32_t *mesh;
    const float *w2o;
    const float *xf;
    const cbvh *blas;
} oscene_t;

/* the object-space ray: o' = W (o, 1), d' = W (d, 0), each row ((w0 x + w1 y) + w2 z) + w3; d' is not
 * normalized, so t is the same parameter in both spaces */
static void object_ray(const float *w, v3 o, v3 d, v3 *oo, v3 *od)
{
    const float ov[3] = {o.x, o.y, o.z}, dv[3] = {d.x, d.y, d.z};
    float ro[3], rd[3];
    for(int a = 0; a < 3; ++a)
    {
        const float *r = w + 4 * a;
        ro[a] = ((r[0] * ov[0] + r[1] * ov[1]) + r[2] * ov[2]) + r[3];
        rd[a] = (r[0] * dv[0] + r[1] * dv[1]) + r[2] * dv[2];
    }
    *oo = V3(ro[0], ro[1], ro[2]);
    *od = V3(rd[0], rd[1], rd[2]);
}

/* rotate_quat (utils/transform.slang:17): v + 2 cross(q.xyz, cross(q.xyz, v) + w v) */
static inline v3 rotq(const float *q, v3 v)
{
    const v3 qv = V3(q[0], q[1], q[2]);
    return add(v, scl(cross(qv, add(cross(qv, v), scl(v, q[3]))), 2.0f));
}

/* apply_transform (utils/transform.slang:19-24): rotate_quat(q, v * scale) + translation */
static inline v3 apply_xf(const float *xf, v3 p)
{
    const v3 r = rotq(xf, V3(p.x * xf[7], p.y * xf[8], p.z * xf[9]));
    return V3(r.x + xf[4], r.y + xf[5], r.z + xf[6]);
}

/* the closest record of a flat scene (is == NULL) or of every instance of a two-level one */
static chit_t scene_closest(const cbvh *flat, const oscene_t *is, v3 o, v3 d, float tmin, float tmax, int cull,
                            int disable_material, float blend_xi)
{
    chit_t h = {tmax, 0.f, 0.f, 0xFFFFFFFFu, 0xFFFFFFFFu};
    if(!is)
    {
        closest(flat, o, d, tmin, 0u, cull, disable_material, blend_xi, &h);
        return h;
    }
    for(uint32_t k = 0; k < is->n_inst; ++k)
    {
        v3 oo, od;
        object_ray(is->w2o + 12 * (size_t)k, o, d, &oo, &od);
        closest(&is->blas[is->mesh[k]], oo, od, tmin, k, cull, disable_material, blend_xi, &h);
    }
    return h;
}

static v3 scene_transmit(const cbvh *flat, const oscene_t *is, const mat_t *mats, const uint32_t *mat_of_prim, v3 o,
                         v3 d, float tmin, float tmax, uint32_t blend_seed, int per_candidate, int cull)
{
    static __thread thit_t hits[ORC_MAX_TRANSMISSIVE];
    uint32_t nh = 0;
    if(!is)
    {
        if(transmit_collect(flat, mats, mat_of_prim, o, d, tmin, tmax, blend_seed, per_candidate, cull, 0u, hits, &nh))
            return V3(0, 0, 0);
    }
    else
        for(uint32_t k = 0; k < is->n_inst; ++k)
        {
            v3 oo, od;
            object_ray(is->w2o + 12 * (size_t)k, o, d, &oo, &od);
            if(transmit_collect(&is->blas[is->mesh[k]], mats, mat_of_prim, oo, od, tmin, tmax, blend_seed,
                                per_candidate, cull, k, hits, &nh))
                return V3(0, 0, 0);
        }
    return transmit_product(mats, mat_of_prim, hits, nh);
}

/* eval_disney (bsdf_disney.slang:298-381), iridescence strength 0 */
static v3 eval_disney(mat_t m, v3 L, v3 N, v3 V, float eta, float *pdf)
{
    int refl = dot(N, L) > 0.0f;
    v3 LpV = refl ? add(L, V) : add(L, scl(V, eta));
    float len = sqrtf(dot(LpV, LpV));
    if(len < 1e-7f)
    {
        *pdf = 0.0f;
        return V3(0, 0, 0);
    }
    v3 H = V3(LpV.x / len, LpV.y / len, LpV.z / len);
    if(dot(V, H) < 0.0f) H = neg(H);
    if(refl && dot(N, H) <= 0.0f)
    {
        *pdf = 0.0f;
        return V3(0, 0, 0);
    }
    float diffuse_ratio = 0.5f * (1.0f - m.metal);
    float psr = 1.0f / (1.0f + m.cc);
    float tw = (1.0f - m.metal) * m.trans;
    m.rough = fmaxf(0.001f, m.rough * m.rough);
    m.sheen_rough = fmaxf(0.001f, m.sheen_rough);
    v3 spec_color = lerp3(V3(0.08f, 0.08f, 0.08f), m.color, m.metal);
    v3 brdf = V3(0, 0, 0), bsdf = V3(0, 0, 0);
    float brdf_pdf = 0.0f, bsdf_pdf = 0.0f, mp = 0.0f;
    if(tw > 0.0f)
    {
        if(refl)
        {
            bsdf = add(bsdf, scl(ev_diel_refl(m.rough, m.color, eta, V, N, L, H, &mp), 1.0f - 0.0f));
            bsdf_pdf += mp * (1.0f - 0.0f);
        }
        else
            bsdf = ev_diel_refr(m.rough, m.color, eta, V, N, L, H, &bsdf_pdf);
    }
    if(tw < 1.0f)
    {
        float dt = m.dt;
        brdf = add(brdf, scl(ev_diffuse(m.rough, m.sheen_rough, m.color, m.sheen, m.metal, V, N, L, H, &mp), 1.0f - dt));
        brdf_pdf += mp * diffuse_ratio * (1.0f - dt);
        if(dt > 0.0f && !refl)
        {
            float cos_t = -dot(N, L);
            brdf = add(brdf, scl(scl(m.dt_color, dt), ONE_OVER_PI));
            brdf_pdf += cos_t * ONE_OVER_PI * diffuse_ratio * dt;
        }
        brdf = add(brdf, scl(ev_spec(m.rough, spec_color, V, N, L, H, &mp), 1.0f - 0.0f));
        brdf_pdf += mp * psr * (1.0f - diffuse_ratio) * (1.0f - 0.0f);
        brdf = add(brdf, ev_cc(m.cc_rough, m.cc, V, N, L, H, &mp));
        brdf_pdf += mp * (1.0f - psr) * (1.0f - diffuse_ratio);
    }
    float p = lerpf(brdf_pdf, bsdf_pdf, tw);
    v3 F = lerp3(brdf, bsdf, tw);
    int valid = p > 0.0f;
    *pdf = valid ? p : 0.0f;
    return valid ? F : V3(0, 0, 0);
}

/* ---- analytic lights: light_t records of 20 words (include/vierkant/punctual_light.hpp:64-89) ---- */
enum { L_OMNI = 0, L_SPOT = 1, L_DIRECTIONAL = 2, L_AREA = 3, L_RECT = 4, L_SPHERE = 5, L_TUBE = 6, L_DISK = 7 };
enum { MIS_FULL = 0, MIS_NEE_ONLY = 1, MIS_BSDF_ONLY = 2 };

typedef struct
{
    v3 position;
    uint32_t type;
    v3 color;
    float intensity;
    v3 direction;
    float range, spot_scale, spot_offset, angular_size, size_y;
    v3 tangent;
    float size_x;
} light_t;

static light_t read_light(const float *w)
{
    light_t l;
    l.position = V3(w[0], w[1], w[2]);
    memcpy(&l.type, &w[3], 4);
    l.color = V3(w[4], w[5], w[6]);
    l.intensity = w[7];
    l.direction = V3(w[8], w[9], w[10]);
    l.range = w[11];
    l.spot_scale = w[12];
    l.spot_offset = w[13];
    l.angular_size = w[14];
    l.size_y = w[15];
    l.tangent = V3(w[16], w[17], w[18]);
    l.size_x = w[19];
    return l;
}

typedef struct
{
    v3 dir;
    float dist, pdf;
    v3 radiance;
} lsample_t;

/* sample_unit_sphere_cap (utils/sampling.slang:55-64) */
static v3 sphere_cap(float xi_x, float xi_y, float sigma)
{
    float phi = 2.0f * PI * xi_y;
    float cs = tcos(sigma);
    float ct = xi_x * (1.0f - cs) + cs;
    float st = sqrtf(1.0f - ct * ct);
    return sph(st, ct, phi);
}

/* sample_light (direct_lighting.slang:32-136) */
static lsample_t sample_light(const light_t *l, v3 P, float xi_x, float xi_y)
{
    lsample_t r;
    r.pdf = 1.0f;
    r.radiance = scl(l->color, l->intensity);
    r.dir = V3(0, 0, 1);
    r.dist = EPS;
    if(l->type == L_DIRECTIONAL)
    {
        r.dir = neg(l->direction);
        r.dist = DEFAULT_TMAX;
        if(l->angular_size > 0.0f)
        {
            r.dir = to_world(local_frame(r.dir), sphere_cap(xi_x, xi_y, l->angular_size));
            r.pdf = 1.0f / (2.0f * PI * (1.0f - tcos(l->angular_size)));
        }
    }
    else if(l->type == L_SPHERE)
    {
        v3 tc = sub(l->position, P);
        float d2 = dot(tc, tc), r2 = l->size_x * l->size_x;
        if(d2 <= r2)
        {
            r.radiance = V3(0, 0, 0);
            r.dir = V3(0, 0, 1);
            r.dist = EPS;
        }
        else
        {
            float s2 = r2 / d2;
            float cs = sqrtf(1.0f - s2);
            r.dir = to_world(local_frame(scl(tc, 1.0f / sqrtf(d2))), sphere_cap(xi_x, xi_y, tasin(sqrtf(s2))));
            r.pdf = (1.0f + cs) / (2.0f * PI * s2);
            float b = dot(r.dir, tc);
            r.dist = b - sqrtf(fmaxf(b * b - d2 + r2, 0.0f));
        }
    }
    else if(l->type == L_RECT || l->type == L_DISK || l->type == L_TUBE)
    {
        v3 Q, n;
        if(l->type == L_RECT)
        {
            v3 bt = cross(l->direction, l->tangent);
            Q = add(add(l->position, scl(l->tangent, (2.0f * xi_x - 1.0f) * l->size_x)),
                    scl(bt, (2.0f * xi_y - 1.0f) * l->size_y));
            n = l->direction;
            r.pdf = 1.0f / (4.0f * l->size_x * l->size_y);
        }
        else if(l->type == L_DISK)
        {
            float rr = sqrtf(xi_x), th = 2.0f * PI * xi_y;
            float dx = rr * tcos(th), dy = rr * tsin(th);
            Q = add(l->position, to_world(local_frame(l->direction), V3(l->size_x * dx, l->size_x * dy, 0.0f)));
            n = l->direction;
            r.pdf = 1.0f / (PI * l->size_x * l->size_x);
        }
        else
        {
            float phi = 2.0f * PI * xi_x;
            n = to_world(local_frame(l->direction), V3(tcos(phi), tsin(phi), 0.0f));
            Q = add(add(l->position, scl(l->direction, (2.0f * xi_y - 1.0f) * l->size_y)), scl(n, l->size_x));
            r.pdf = 1.0f / (4.0f * PI * l->size_x * l->size_y);
        }
        v3 tq = sub(Q, P);
        float d2 = dot(tq, tq);
        r.dist = fmaxf(sqrtf(d2), EPS);
        r.dir = V3(tq.x / r.dist, tq.y / r.dist, tq.z / r.dist);
        float cl = dot(n, neg(r.dir));
        r.pdf *= d2 / fmaxf(cl, EPS);
        if(cl <= 0.0f) r.radiance = V3(0, 0, 0);
    }
    else
    {
        v3 tl = sub(l->position, P);
        r.dist = fmaxf(sqrtf(dot(tl, tl)), EPS);
        r.dir = V3(tl.x / r.dist, tl.y / r.dist, tl.z / r.dist);
        float w = r.dist / l->range;
        w *= w;
        float att = clampf(1.0f - w * w, 0.0f, 1.0f) / (r.dist * r.dist);
        if(l->type == L_SPOT)
        {
            float cd = dot(l->direction, neg(r.dir));
            float a = clampf(cd * l->spot_scale + l->spot_offset, 0.0f, 1.0f);
            att *= a * a;
        }
        r.radiance = scl(r.radiance, att);
    }
    return r;
}

static float power_heuristic(float pa, float pb)
{
    float r = pb / pa;
    return 1.0f / (1.0f + r * r);
}

static int bsdf_reachable(const light_t *l)
{
    return (l->type == L_DIRECTIONAL && l->angular_size > 0.0f) || l->type == L_RECT || l->type == L_SPHERE ||
           l->type == L_TUBE || l->type == L_DISK;
}

/* intersect_light (direct_lighting.slang:166-236): front faces of light bodies; t <= 0 is a miss */
static float intersect_light(const light_t *l, v3 P, v3 d, float *cos_light)
{
    *cos_light = 0.0f;
    if(l->type == L_SPHERE)
    {
        v3 tc = sub(l->position, P);
        float d2 = dot(tc, tc), r2 = l->size_x * l->size_x;
        if(d2 <= r2) return 0.0f;
        float b = dot(d, tc);
        float disc = b * b - d2 + r2;
        if(disc <= 0.0f) return 0.0f;
        float t = b - sqrtf(disc);
        if(t <= 0.0f) return 0.0f;
        v3 q = sub(add(P, scl(d, t)), l->position);
        *cos_light = dot(V3(q.x / l->size_x, q.y / l->size_x, q.z / l->size_x), neg(d));
        return t;
    }
    if(l->type == L_RECT || l->type == L_DISK)
    {
        float cl = dot(l->direction, neg(d));
        if(cl <= 0.0f) return 0.0f;
        float t = dot(sub(l->position, P), l->direction) / -cl;
        if(t <= 0.0f) return 0.0f;
        v3 q = sub(add(P, scl(d, t)), l->position);
        if(l->type == L_RECT)
        {
            v3 bt = cross(l->direction, l->tangent);
            if(fabsf(dot(q, l->tangent)) > l->size_x || fabsf(dot(q, bt)) > l->size_y) return 0.0f;
        }
        else if(dot(q, q) > l->size_x * l->size_x)
            return 0.0f;
        *cos_light = cl;
        return t;
    }
    if(l->type == L_TUBE)
    {
        v3 o = sub(P, l->position);
        v3 dp = sub(d, scl(l->direction, dot(d, l->direction)));
        v3 op = sub(o, scl(l->direction, dot(o, l->direction)));
        float A = dot(dp, dp), B = dot(op, dp), C = dot(op, op) - l->size_x * l->size_x;
        float disc = B * B - A * C;
        if(A <= EPS || disc <= 0.0f || C <= 0.0f) return 0.0f;
        float t = (-B - sqrtf(disc)) / A;
        if(t <= 0.0f || fabsf(dot(add(o, scl(d, t)), l->direction)) > l->size_y) return 0.0f;
        v3 nq = add(op, scl(dp, t));
        float cl = dot(V3(nq.x / l->size_x, nq.y / l->size_x, nq.z / l->size_x), neg(d));
        if(cl <= 0.0f) return 0.0f;
        *cos_light = cl;
        return t;
    }
    return 0.0f;
}

static float mis_weight_nee(uint32_t mode, float p_nee, float p_bsdf)
{
    if(mode == MIS_NEE_ONLY) return 1.0f;
    if(mode == MIS_BSDF_ONLY) return 0.0f;
    return power_heuristic(p_nee, p_bsdf);
}

static float mis_weight_hit(uint32_t mode, float prev_pdf, float p_nee)
{
    if(mode == MIS_NEE_ONLY) return 0.0f;
    if(mode == MIS_BSDF_ONLY || prev_pdf <= 0.0f) return 1.0f;
    return power_heuristic(prev_pdf, p_nee);
}

/* light_pdf (direct_lighting.slang:254-292) */
static float light_pdf(const light_t *l, v3 P, v3 d)
{
    if(l->type == L_DIRECTIONAL)
    {
        if(l->angular_size <= 0.0f) return 0.0f;
        float cs = tcos(l->angular_size);
        return dot(d, neg(l->direction)) < cs ? 0.0f : 1.0f / (2.0f * PI * (1.0f - cs));
    }
    if(l->type == L_SPHERE)
    {
        v3 tc = sub(l->position, P);
        float d2 = dot(tc, tc), r2 = l->size_x * l->size_x;
        if(d2 <= r2) return 0.0f;
        float s2 = r2 / d2;
        float cs = sqrtf(1.0f - s2);
        return dot(d, scl(tc, 1.0f / sqrtf(d2))) < cs ? 0.0f : (1.0f + cs) / (2.0f * PI * s2);
    }
    if(l->type == L_RECT || l->type == L_DISK || l->type == L_TUBE)
    {
        float cl;
        float t = intersect_light(l, P, d, &cl);
        if(t <= 0.0f) return 0.0f;
        float area = l->type == L_RECT ? 4.0f * l->size_x * l->size_y
                     : l->type == L_DISK ? PI * l->size_x * l->size_x
                                         : 4.0f * PI * l->size_x * l->size_y;
        return t * t / (area * cl);
    }
    return 0.0f;
}

/* ---- the integrator ------------------------------------------------------------------------- */
typedef struct
{
    const orc_pt_scene *sc;
    const orc_pt_params *pp;
    cbvh bvh;
    mat_t *mats;
    const uint32_t *mat_of_prim;
    const light_t *lights;
    const otex_t *tex;
    const float *uv;  /* 2 per vertex */
    const float *tan; /* 3 per vertex */
    float *obj_tmax;
    const float *obj_he;   /* per object: the half extents of its box (3 floats) */
    const oscene_t *inst;  /* NULL for a flat scene */
    float *out; /* 8 floats per pixel: r g b alpha depth normal(bits) object(bits) albedo(bits) */
    uint32_t y0, y1;
} job_t;

/* the segment's medium, absorption only (raypipeline.slang:380-415 with sigma_s = 0): a channel picked, a
 * collision distance drawn against its extinction; a collision before t_hit absorbs the path (returns 1),
 * else the throughput takes exp(-sigma t_hit) over the mean of the three channels' transmittances */
static int medium_collision(const mat_t *m, float t_hit, v3 *beta, uint32_t *rng)
{
    const v3 sg = m->sigma;
    if(!(sg.x > 0.f || sg.y > 0.f || sg.z > 0.f)) return 0;
    const int channel = (int)fminf(rnd(rng) * 3.0f, 2.0f);
    const float sc = channel == 0 ? sg.x : (channel == 1 ? sg.y : sg.z);
    const float t = fminf(-logf(1.0f - rnd(rng)) / fmaxf(sc, EPS), t_hit);
    if(t < t_hit) return 1;
    const v3 tr = V3(expf(-sg.x * t), expf(-sg.y * t), expf(-sg.z * t));
    const float pdf = fmaxf((tr.x + tr.y + tr.z) / 3.0f, PDF_EPS);
    *beta = V3(beta->x * (tr.x / pdf), beta->y * (tr.y / pdf), beta->z * (tr.z / pdf));
    return 0;
}

static void render_pixel(job_t *j, uint32_t x, uint32_t y)
{
    const orc_pt_scene *sc = j->sc;
    const orc_pt_params *pp = j->pp;
    const uint32_t W = pp->width, H = pp->height;
    uint32_t rng = orc_xxhash32(pp->seed, W * y + x);
    /* the ray cone's spread angle (raypipeline.slang:102), a half like the payload's packed cone */
    const float cone_spread = orc_half_value(tatan(2.0f * ttan(pp->fov_y / 2.0f) / (float)H));
    v3 result = V3(0, 0, 0);
    float coverage = 0.f;
    float offset = 2.0f * rnd(&rng) - 1.0f;
    v3 eye = V3(pp->eye[0], pp->eye[1], pp->eye[2]);
    v3 fwd = V3(pp->forward[0], pp->forward[1], pp->forward[2]);
    v3 right = V3(pp->right[0], pp->right[1], pp->right[2]);
    v3 up = V3(pp->up[0], pp->up[1], pp->up[2]);
    float zdepth = 0.f;
    uint32_t first_obj = 0xFFFFu;
    v3 first_n = V3(0, 0, 0), first_alb = V3(0, 0, 0);
    for(uint32_t s = 0; s < pp->num_samples; ++s)
    {
        float hx = (float)s / (float)pp->num_samples;
        float hy = (float)brev(s) * 2.3283064365386963e-10f;
        float fx = hx + offset, fy = hy + offset;
        float cx = (float)x + (fx - floorf(fx));
        float cy = (float)y + (fy - floorf(fy));
        float xi_x = rnd(&rng), xi_y = rnd(&rng);
        /* camera ray */
        float u = cx / (float)W, v = cy / (float)H;
        float dx = u * 2.0f - 1.0f, dy = v * 2.0f - 1.0f;
        float r = sqrtf(xi_x), th = 2.0f * PI * xi_y;
        float lx = pp->aperture / 2.0f * (r * tcos(th)), ly = pp->aperture / 2.0f * (r * tsin(th));
        v3 org = add(add(eye, scl(right, lx)), scl(up, ly));
        v3 tg = nrmz(V3(dx * pp->tan_half_x, -dy * pp->tan_half_y, -1.0f));
        v3 fp = scl(tg, pp->focal_distance / fabsf(tg.z));
        v3 dv = sub(fp, V3(lx, ly, 0.f));
        v3 dir = nrmz(sub(add(scl(right, dv.x), scl(up, dv.y)), scl(fwd, dv.z)));
        float tmax = DEFAULT_TMAX;
        uint32_t media_index = 0;
        float iors[4] = {1.f, 1.f, 1.f, 1.f};
        /* the material whose interior each stack level is (level 0: vacuum) */
        uint32_t media_mat[4] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
        int regularize = 0, caustic = 0, first_miss = 1;
        v3 beta = V3(1, 1, 1), rad = V3(0, 0, 0);
        float prev_pdf = 0.0f; /* payload.prev_bsdf_pdf */
        float cone_width = 0.0f;
        const uint32_t nl = pp->n_lights;
        for(uint32_t i = 0; i < pp->max_depth; ++i)
        {
            rng = orc_xxhash32(rng, i);
            uint32_t lr = rng;
            const v3 seg_o = org, seg_d = dir, seg_beta = beta;
            const float seg_tmax = tmax, seg_prev = prev_pdf;
            const int seg_caustic = caustic;
            float t = 0.f, bu = 0.f, bv = 0.f;
            /* BLEND candidates compare against rnd() of a copy of payload.rng_state (raypipeline.slang:693) */
            uint32_t peek = rng;
            const float blend_xi = rnd(&peek);
            const chit_t ch = scene_closest(&j->bvh, j->inst, org, dir, EPS, tmax, media_index == 0,
                                            (int)pp->disable_material, blend_xi);
            const uint32_t prim = ch.prim;
            t = ch.t;
            bu = ch.u;
            bv = ch.v;
            int stop = 0, media_op = 0;
            float next_ior = 1.f;
            uint32_t obj = 0xFFFFu;
            v3 pos = V3(0, 0, 0), nrm = V3(0, 0, 0), alb = V3(0, 0, 0);
            if(prim == 0xFFFFFFFFu)
            {
                stop = 1;
                rad = add(rad, mul(scl(beta, sc->environment), env_white(dir)));
                if(nl && !(pp->suppress_refractive_caustics && caustic))
                {
                    /* eval_distant_lights (raypipeline.slang:728-744) */
                    v3 dl = V3(0, 0, 0);
                    for(uint32_t li = 0; li < nl; ++li)
                    {
                        const light_t *l = &j->lights[li];
                        if(l->type != L_DIRECTIONAL) continue;
                        float pdf = light_pdf(l, V3(0, 0, 0), dir);
                        if(pdf <= 0.0f) continue;
                        float w = mis_weight_hit(pp->mis_mode, prev_pdf, pdf / (float)nl);
                        dl = add(dl, scl(scl(l->color, w), l->intensity));
                    }
                    rad = add(rad, mul(beta, dl));
                }
            }
            else if(media_mat[media_index] != 0xFFFFFFFFu && medium_collision(&j->mats[media_mat[media_index]], t,
                                                                            &beta, &lr))
            {
                /* absorbed in the medium before the surface (raypipeline.slang:380-415 with a single-scattering
                 * albedo of 0: the throughput becomes 0 there, so the path gathers nothing more) */
                stop = 1;
            }
            else
            {
                const uint32_t *ix = sc->indices + 3 * (size_t)prim;
                const float *P = sc->positions, *N = sc->normals;
                v3 v0 = V3(P[3 * ix[0]], P[3 * ix[0] + 1], P[3 * ix[0] + 2]);
                v3 v1 = V3(P[3 * ix[1]], P[3 * ix[1] + 1], P[3 * ix[1] + 2]);
                v3 v2 = V3(P[3 * ix[2]], P[3 * ix[2] + 1], P[3 * ix[2] + 2]);
                v3 e1 = sub(v1, v0), e2 = sub(v2, v0);
                /* HitKind from the determinant in the triangle's own space; an instance's shading point and
                 * normal go through its transform (interpolate_vertex, raypipeline.slang:336-358) */
                const float *xf = j->inst ? j->inst->xf + 10 * (size_t)ch.inst : NULL;
                v3 dir_o = dir;
                if(j->inst)
                {
                    v3 unused_o;
                    object_ray(j->inst->w2o + 12 * (size_t)ch.inst, V3(0, 0, 0), dir, &unused_o, &dir_o);
                }
                int backface = dot(e1, cross(dir_o, e2)) < 0.f;
                float bx = 1.0f - bu - bv, by = bu, bz = bv;
                pos = add(add(scl(v0, bx), scl(v1, by)), scl(v2, bz));
                v3 n = add(add(scl(V3(N[3 * ix[0]], N[3 * ix[0] + 1], N[3 * ix[0] + 2]), bx),
                               scl(V3(N[3 * ix[1]], N[3 * ix[1] + 1], N[3 * ix[1] + 2]), by)),
                           scl(V3(N[3 * ix[2]], N[3 * ix[2] + 1], N[3 * ix[2] + 2]), bz));
                nrm = nrmz(n);
                if(xf)
                {
                    pos = apply_xf(xf, pos);
                    nrm = rotq(xf, nrm);
                }
                if(backface) nrm = neg(nrm);
                obj = sc->tri_object[prim];
                /* tmax_inside_volume = 2 * length(scale * aabb.half_extents()) (raypipeline.slang:370) */
                float tmax_inside = j->obj_tmax[obj];
                if(xf)
                {
                    const float *he = j->obj_he + 3 * (size_t)obj;
                    const float hx = xf[7] * he[0], hy = xf[8] * he[1], hz = xf[9] * he[2];
                    tmax_inside = 2.f * sqrtf((hx * hx + hy * hy) + hz * hz);
                }
                mat_t m = j->mats[sc->object_material[obj]];
                v3 Vw = neg(dir);
                /* before the transmission texture, or an interior medium (raypipeline.slang:517-522) */
                const int media_boundary = m.trans > 0.0f || m.vol;
                if(m.tex | m.tex_normal | m.tex_emission | m.tex_arm | m.tex_transmission | m.tex_dt | m.tex_dtc)
                {
                    /* every texture at the ray cone's level of detail (raypipeline.slang:300-321, 368-375,
                     * 526-606): 0.5 log2(uv area / world area) + log2|cone width| + 0.5 log2(w h) - log2(N.V),
                     * with the width the cone had before this segment */
                    const float *T0 = j->uv + 2 * (size_t)ix[0], *T1 = j->uv + 2 * (size_t)ix[1],
                                *T2 = j->uv + 2 * (size_t)ix[2];
                    float tu = (T0[0] * bx + T1[0] * by) + T2[0] * bz, tv = (T0[1] * bx + T1[1] * by) + T2[1] * bz;
                    /* lod_constant: the triangle's area in world space (raypipeline.slang:299-309) */
                    const v3 w0 = xf ? apply_xf(xf, v0) : v0, w1 = xf ? apply_xf(xf, v1) : v1,
                             w2 = xf ? apply_xf(xf, v2) : v2;
                    v3 pc = cross(sub(w1, w0), sub(w2, w0));
                    float p_a = sqrtf(dot(pc, pc));
                    float t_a = fabsf((T1[0] - T0[0]) * (T2[1] - T0[1]) - (T2[0] - T0[0]) * (T1[1] - T0[1]));
                    const float lod_cone = 0.5f * tlog2(t_a / p_a) + tlog2(fabsf(cone_width));
                    const float lod_nov = tlog2(fabsf(dot(Vw, nrm)));
                    float c[4];
#define ORC_LOOKUP(slot)                                                                                   \
    do                                                                                                     \
    {                                                                                                      \
        const otex_t *tx = &j->tex[(slot) - 1];                                                           \
        float lambda = lod_cone;                                                                           \
        lambda += 0.5f * tlog2((float)tx->w * (float)tx->h);                                              \
        lambda -= lod_nov;                                                                                 \
        otex_sample(tx, tu, tv, lambda, c);                                                                \
    } while(0)
                    if(m.tex)
                    {
                        ORC_LOOKUP(m.tex);
                        m.color = V3(m.color.x * c[0], m.color.y * c[1], m.color.z * c[2]);
                        m.alpha = m.alpha * c[3];
                    }
                    if(m.tex_normal)
                    {
                        /* tangent-space normal, bitangent flipped on back faces, hemisphere clamp (:536-552) */
                        ORC_LOOKUP(m.tex_normal);
                        v3 nt;
                        nt.x = 2.0f * (c[0] - 0.5f);
                        nt.y = 2.0f * (c[1] - 0.5f);
                        nt.z = sqrtf(fmaxf(0.0f, 1.0f - nt.x * nt.x - nt.y * nt.y));
                        const float *G = j->tan;
                        v3 tg = nrmz(add(add(scl(V3(G[3 * ix[0]], G[3 * ix[0] + 1], G[3 * ix[0] + 2]), bx),
                                             scl(V3(G[3 * ix[1]], G[3 * ix[1] + 1], G[3 * ix[1] + 2]), by)),
                                         scl(V3(G[3 * ix[2]], G[3 * ix[2] + 1], G[3 * ix[2] + 2]), bz)));
                        if(xf) tg = rotq(xf, tg);
                        v3 bt = nrmz(cross(nrm, tg));
                        if(backface) bt = neg(bt);
                        v3 nm = nrmz(add(add(scl(tg, nt.x), scl(bt, nt.y)), scl(nrm, nt.z)));
                        nrm = nrmz(add(nm, scl(nrm, fmaxf(0.0f, -dot(nm, nrm) + 1e-3f))));
                    }
                    if(m.tex_emission)
                    {
                        ORC_LOOKUP(m.tex_emission);
                        if(m.emission.x > 0.0f || m.emission.y > 0.0f || m.emission.z > 0.0f)
                            m.emission = V3(m.emission.x * c[0], m.emission.y * c[1], m.emission.z * c[2]);
                        else
                            m.emission = V3(c[0], c[1], c[2]);
                    }
                    if(m.tex_arm)
                    {
                        ORC_LOOKUP(m.tex_arm);
                        m.rough = m.rough * c[1];
                        m.metal = m.metal * c[2];
                    }
                    if(m.tex_transmission)
                    {
                        ORC_LOOKUP(m.tex_transmission);
                        m.trans = m.trans * c[0];
                    }
                    if(m.tex_dt)
                    {
                        ORC_LOOKUP(m.tex_dt);
                        m.dt = m.dt * c[3];
                    }
                    if(m.tex_dtc)
                    {
                        ORC_LOOKUP(m.tex_dtc);
                        m.dt_color = V3(m.dt_color.x * c[0], m.dt_color.y * c[1], m.dt_color.z * c[2]);
                    }
#undef ORC_LOOKUP
                }
                /* propagate(cone, 0, hit t) (raypipeline.slang:280-286, :608); the payload keeps halves */
                cone_width = orc_half_value(cone_width + cone_spread * t);
                if(pp->disable_material) m.color = V3(0.8f, 0.8f, 0.8f);
                org = pos;
                float eta = backface ? m.ior / iors[media_index > 0 ? media_index - 1 : 0] : iors[media_index] / m.ior;
                eta += EPS;
                next_ior = m.ior;
                media_op = !media_boundary ? 0 : (backface ? 2 : 1);
                if(i == 0) alb = m.color;
                v3 ff = dot(dir, nrm) < 0.f ? nrm : neg(nrm);
                v3 em = scl(m.emission, dot(nrm, ff) > 0.f ? m.emission_strength : 0.0f);
                rad = add(rad, mul(beta, em));
                if(regularize) m.rough = fmaxf(m.rough, 0.25f);
                bs_t b = sample_disney(m, nrm, Vw, eta, &lr);
                if(nl)
                {
                    /* nee_surface (direct_lighting.slang:296-336) */
                    uint32_t index = 0;
                    if(nl > 1)
                    {
                        uint32_t k = (uint32_t)(rnd(&lr) * (float)nl);
                        index = k < nl - 1 ? k : nl - 1;
                    }
                    float xi_x = rnd(&lr), xi_y = rnd(&lr);
                    const light_t *l = &j->lights[index];
                    lsample_t ls = sample_light(l, pos, xi_x, xi_y);
                    if(ls.radiance.x > 0.0f || ls.radiance.y > 0.0f || ls.radiance.z > 0.0f)
                    {
                        v3 so = add(pos, scl(nrm, EPS));
                        v3 T = scene_transmit(&j->bvh, j->inst, j->mats, j->mat_of_prim, so, ls.dir, 0.0f, ls.dist,
                                              lr, 1, 0);
                        if(T.x > 0.0f || T.y > 0.0f || T.z > 0.0f)
                        {
                            float bpdf = 0.0f;
                            float ct = fmaxf(dot(nrm, ls.dir), 0.0f);
                            v3 F = eval_disney(m, ls.dir, nrm, Vw, eta, &bpdf);
                            float w = bsdf_reachable(l) ? mis_weight_nee(pp->mis_mode, ls.pdf / (float)nl, bpdf) : 1.0f;
                            v3 nee = scl(T, w);
                            nee = scl(nee, (float)nl);
                            nee = mul(nee, ls.radiance);
                            nee = mul(nee, F);
                            nee = scl(nee, ct);
                            nee = V3(nee.x / ls.pdf, nee.y / ls.pdf, nee.z / ls.pdf);
                            rad = add(rad, mul(beta, nee));
                        }
                    }
                }
                if(!(b.pdf > 0.0f)) stop = 1;
                else
                {
                    prev_pdf = nl ? b.pdf : 0.0f;
                    dir = b.dir;
                    int specular = b.pdf < 1e3f;
                    regularize = regularize || specular;
                    caustic = specular && (caustic || b.transmission);
                    float ct = fabsf(dot(nrm, b.dir));
                    float den = fmaxf(b.pdf, PDF_EPS);
                    v3 thr = scl(b.F, ct);
                    thr = V3(thr.x / den, thr.y / den, thr.z / den);
                    if(i >= 1 && pp->max_path_beta > 0.0f)
                        thr = V3(fminf(thr.x, pp->max_path_beta), fminf(thr.y, pp->max_path_beta),
                                 fminf(thr.z, pp->max_path_beta));
                    beta = mul(beta, thr);
                    if(b.transmission) tmax = media_op == 1 ? tmax_inside : DEFAULT_TMAX;
                    media_op = b.transmission ? media_op : 0;
                    if(!(pp->test_flags & ORC_PT_NO_ROULETTE) && max3(beta) <= 0.05f * (float)i)
                    {
                        float q = 1.0f - max3(beta);
                        if(rnd(&lr) < q) stop = 1;
                        else
                        {
                            float keep = 1.0f - q;
                            beta = V3(beta.x / keep, beta.y / keep, beta.z / keep);
                        }
                    }
                }
            }
            if(media_op == 1)
            {
                media_index = media_index + 1 < 3 ? media_index + 1 : 3;
                iors[media_index] = round_half(next_ior);
                media_mat[media_index] = sc->object_material[obj];
            }
            else if(media_op == 2)
                media_index = media_index > 0 ? media_index - 1 : 0;
            if(nl)
            {
                /* ghost-light pass over the traced segment (raypipeline.slang:201-222) */
                v3 dv = sub(pos, seg_o);
                float t_end = prim == 0xFFFFFFFFu ? seg_tmax : sqrtf(dot(dv, dv));
                for(uint32_t li = 0; li < nl; ++li)
                {
                    const light_t *l = &j->lights[li];
                    float cl;
                    float th = intersect_light(l, seg_o, seg_d, &cl);
                    if(th <= EPS || th >= t_end) continue;
                    if(pp->suppress_refractive_caustics && seg_caustic) continue;
                    float p_nee = light_pdf(l, seg_o, seg_d) / (float)nl;
                    float w = mis_weight_hit(pp->mis_mode, seg_prev, p_nee);
                    rad = add(rad, scl(mul(scl(seg_beta, w), l->color), l->intensity));
                }
            }
            if(i == 0)
            {
                first_miss = obj == 0xFFFFu;
                first_obj = obj;
                first_n = nrm;
                first_alb = alb;
                zdepth = first_miss ? 0.f : dot(sub(pos, eye), fwd);
            }
            if(stop) break;
        }
        result = add(result, rad);
        coverage = fmaxf(coverage, first_miss ? 0.f : 1.f);
    }
    float ns = (float)pp->num_samples;
    float *o = j->out + 8 * ((size_t)y * W + x);
    o[0] = result.x / ns;
    o[1] = result.y / ns;
    o[2] = result.z / ns;
    o[3] = coverage;
    o[4] = zdepth;
    (void)first_n;
    (void)first_alb;
    memcpy(&o[6], &first_obj, 4);
    o[5] = 0.f;
    o[7] = 0.f;
}

static void *render_rows(void *arg)
{
    job_t *j = (job_t *)arg;
    for(uint32_t y = j->y0; y < j->y1; ++y)
        for(uint32_t x = 0; x < j->pp->width; ++x) render_pixel(j, x, y);
    return NULL;
}

void orc_sample_disney_n(const float *mat36, const float *N, const float *V, float eta, uint32_t seed, uint32_t n,
                         float *out8)
{
    const mat_t m = unpack_material(mat36, 0);
    const v3 nn = V3(N[0], N[1], N[2]), vv = V3(V[0], V[1], V[2]);
    for(uint32_t i = 0; i < n; ++i)
    {
        uint32_t rng = orc_xxhash32(seed, i);
        const bs_t b = sample_disney(m, nn, vv, eta, &rng);
        float *o = out8 + 8 * (size_t)i;
        o[0] = b.dir.x, o[1] = b.dir.y, o[2] = b.dir.z, o[3] = b.pdf;
        o[4] = b.F.x, o[5] = b.F.y, o[6] = b.F.z, o[7] = (float)b.transmission;
    }
}

void orc_eval_disney_n(const float *mat36, const float *N, const float *V, float eta, const float *L, uint32_t n,
                       float *out4)
{
    const mat_t m = unpack_material(mat36, 0);
    const v3 nn = V3(N[0], N[1], N[2]), vv = V3(V[0], V[1], V[2]);
    for(uint32_t i = 0; i < n; ++i)
    {
        float pdf = 0.f;
        const v3 F = eval_disney(m, V3(L[3 * i], L[3 * i + 1], L[3 * i + 2]), nn, vv, eta, &pdf);
        float *o = out4 + 4 * (size_t)i;
        o[0] = F.x, o[1] = F.y, o[2] = F.z, o[3] = pdf;
    }
}

/* everything a render or a transmittance query reads, built once per call from an orc_pt_scene */
typedef struct
{
    cbvh flat;    /* the flat scene's BVH (n_instances == 0) */
    cbvh *blas;   /* one per mesh (n_instances > 0) */
    oscene_t is;
    float *tris;
    uint8_t *two;
    mat_t *mats;
    uint32_t *mat_of_prim;
    float *obj_tmax, *obj_he;
    float *uv, *tan;
    otex_t *tex;
    alpha_ctx ac;
} built_t;

static void bvh_over(cbvh *b, const float *tris, const uint8_t *two, const double *cent, uint32_t first, uint32_t n)
{
    memset(b, 0, sizeof(*b));
    b->tris = tris;
    b->two_sided = two;
    b->n = n;
    b->order = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    for(uint32_t t = 0; t < n; ++t) b->order[t] = first + t;
    b->nodes = (cnode *)malloc(sizeof(cnode) * (2 * (size_t)(n ? n : 1)));
    if(n) build_rec(b, 0, n, cent);
}

static void scene_build(const orc_pt_scene *sc, built_t *s)
{
    memset(s, 0, sizeof(*s));
    const uint32_t n = sc->n_tri;
    s->tris = (float *)malloc(sizeof(float) * 9 * (size_t)(n ? n : 1));
    s->two = (uint8_t *)calloc(n ? n : 1, 1);
    double *cent = (double *)malloc(sizeof(double) * 3 * (size_t)(n ? n : 1));
    s->mats = (mat_t *)calloc(sc->n_materials ? sc->n_materials : 1, sizeof(mat_t));
    for(uint32_t m = 0; m < sc->n_materials; ++m)
        s->mats[m] = unpack_material(sc->materials + 40 * (size_t)m, sc->n_textures);
    for(uint32_t t = 0; t < n; ++t)
    {
        for(int k = 0; k < 3; ++k)
            for(int a = 0; a < 3; ++a)
                s->tris[9 * (size_t)t + 3 * k + a] = sc->positions[3 * (size_t)sc->indices[3 * (size_t)t + k] + a];
        s->two[t] = (uint8_t)s->mats[sc->object_material[sc->tri_object[t]]].two_sided;
        const float *tr = s->tris + 9 * (size_t)t;
        for(int a = 0; a < 3; ++a) cent[3 * (size_t)t + a] = ((double)tr[a] + tr[3 + a] + tr[6 + a]) / 3.0;
    }
    if(sc->n_instances)
    {
        s->blas = (cbvh *)calloc(sc->n_meshes ? sc->n_meshes : 1, sizeof(cbvh));
        for(uint32_t m = 0; m < sc->n_meshes; ++m)
            bvh_over(&s->blas[m], s->tris, s->two, cent, sc->mesh_first[m], sc->mesh_first[m + 1] - sc->mesh_first[m]);
        s->is.n_inst = sc->n_instances;
        s->is.mesh = sc->inst_mesh;
        s->is.w2o = sc->inst_w2o;
        s->is.xf = sc->inst_xf;
        s->is.blas = s->blas;
    }
    else
        bvh_over(&s->flat, s->tris, s->two, cent, 0, n);
    free(cent);
    /* per object: the half extents of its box and tmax inside a volume, 2 * length(half extents) */
    const uint32_t no = sc->n_objects ? sc->n_objects : 1;
    s->obj_tmax = (float *)calloc(no, sizeof(float));
    s->obj_he = (float *)calloc(3 * (size_t)no, sizeof(float));
    float *lo = (float *)malloc(sizeof(float) * 3 * no), *hi = (float *)malloc(sizeof(float) * 3 * no);
    for(uint32_t o = 0; o < 3 * sc->n_objects; ++o)
    {
        lo[o] = FLT_MAX;
        hi[o] = -FLT_MAX;
    }
    for(uint32_t t = 0; t < n; ++t)
    {
        uint32_t o = sc->tri_object[t];
        for(int k = 0; k < 3; ++k)
            for(int a = 0; a < 3; ++a)
            {
                float c = s->tris[9 * (size_t)t + 3 * k + a];
                lo[3 * o + a] = c < lo[3 * o + a] ? c : lo[3 * o + a];
                hi[3 * o + a] = hi[3 * o + a] < c ? c : hi[3 * o + a];
            }
    }
    for(uint32_t o = 0; o < sc->n_objects; ++o)
    {
        float *he = s->obj_he + 3 * (size_t)o;
        for(int a = 0; a < 3; ++a) he[a] = hi[3 * o] < lo[3 * o] ? 0.f : (hi[3 * o + a] - lo[3 * o + a]) / 2.f;
        s->obj_tmax[o] = 2.f * sqrtf((he[0] * he[0] + he[1] * he[1]) + he[2] * he[2]);
    }
    free(lo);
    free(hi);
    s->mat_of_prim = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    for(uint32_t t = 0; t < n; ++t) s->mat_of_prim[t] = sc->object_material[sc->tri_object[t]];
    s->uv = (float *)calloc(2 * (size_t)(sc->n_vert ? sc->n_vert : 1), sizeof(float));
    if(sc->texcoords) memcpy(s->uv, sc->texcoords, sizeof(float) * 2 * (size_t)sc->n_vert);
    s->tan = (float *)calloc(3 * (size_t)(sc->n_vert ? sc->n_vert : 1), sizeof(float));
    if(sc->tangents) memcpy(s->tan, sc->tangents, sizeof(float) * 3 * (size_t)sc->n_vert);
    s->tex = (otex_t *)calloc(sc->n_textures ? sc->n_textures : 1, sizeof(otex_t));
    for(uint32_t k = 0; k < sc->n_textures; ++k) otex_build(&s->tex[k], &sc->textures[k]);
    s->ac = (alpha_ctx){s->mats, s->mat_of_prim, sc->indices, s->uv, s->tex};
    /* the BVHs point at ac: scene_build's caller keeps `s` in place until scene_free */
    s->flat.ac = &s->ac;
    for(uint32_t m = 0; m < sc->n_meshes && s->blas; ++m) s->blas[m].ac = &s->ac;
}

static void scene_free(const orc_pt_scene *sc, built_t *s)
{
    for(uint32_t k = 0; k < sc->n_textures; ++k) otex_free(&s->tex[k]);
    free(s->tex);
    free(s->uv);
    free(s->tan);
    free(s->mat_of_prim);
    free(s->obj_tmax);
    free(s->obj_he);
    free(s->flat.nodes);
    free(s->flat.order);
    for(uint32_t m = 0; m < sc->n_meshes && s->blas; ++m)
    {
        free(s->blas[m].nodes);
        free(s->blas[m].order);
    }
    free(s->blas);
    free(s->mats);
    free(s->two);
    free(s->tris);
}

void orc_pt_render(const orc_pt_scene *sc, const orc_pt_params *pp, float *out8, int nthreads)
{
    g_ulp_bias = (pp->test_flags & ORC_PT_ULP_BIAS) != 0;
    built_t s;
    scene_build(sc, &s);
    light_t *lights = (light_t *)calloc(pp->n_lights ? pp->n_lights : 1, sizeof(light_t));
    for(uint32_t li = 0; li < pp->n_lights; ++li) lights[li] = read_light(pp->lights + 20 * (size_t)li);
    if(nthreads < 1) nthreads = 1;
    pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
    job_t *jobs = (job_t *)calloc((size_t)nthreads, sizeof(job_t));
    for(int k = 0; k < nthreads; ++k)
    {
        jobs[k].sc = sc;
        jobs[k].pp = pp;
        jobs[k].bvh = s.flat;
        jobs[k].mats = s.mats;
        jobs[k].mat_of_prim = s.mat_of_prim;
        jobs[k].lights = lights;
        jobs[k].tex = s.tex;
        jobs[k].uv = s.uv;
        jobs[k].tan = s.tan;
        jobs[k].obj_tmax = s.obj_tmax;
        jobs[k].obj_he = s.obj_he;
        jobs[k].inst = sc->n_instances ? &s.is : NULL;
        jobs[k].out = out8;
        jobs[k].y0 = (uint32_t)((uint64_t)pp->height * k / nthreads);
        jobs[k].y1 = (uint32_t)((uint64_t)pp->height * (k + 1) / nthreads);
        pthread_create(&th[k], NULL, render_rows, &jobs[k]);
    }
    for(int k = 0; k < nthreads; ++k) pthread_join(th[k], NULL);
    free(th);
    free(jobs);
    free(lights);
    scene_free(sc, &s);
    g_ulp_bias = 0;
}

/* vkrt_pt_trace_transmittance's query on the same scene: per ray (8 floats: origin, tmin, direction, tmax)
 * the transmittance_test product, with BLEND candidates kept (threshold 0) and MASK ones cut at their cutoff,
 * back faces of one-sided triangles skipped when `cull`. out4: rgb, 0 per ray. */
void orc_pt_transmittance(const orc_pt_scene *sc, const float *rays8, uint32_t n, int cull, float *out4)
{
    built_t s;
    scene_build(sc, &s);
    const oscene_t *is = sc->n_instances ? &s.is : NULL;
    for(uint32_t r = 0; r < n; ++r)
    {
        const float *ray = rays8 + 8 * (size_t)r;
        const v3 T = scene_transmit(&s.flat, is, s.mats, s.mat_of_prim, V3(ray[0], ray[1], ray[2]),
                                    V3(ray[4], ray[5], ray[6]), ray[3], ray[7], 0u, 0, cull);
        out4[4 * (size_t)r] = T.x;
        out4[4 * (size_t)r + 1] = T.y;
        out4[4 * (size_t)r + 2] = T.z;
        out4[4 * (size_t)r + 3] = 0.f;
    }
    scene_free(sc, &s);
}

/* ------------------------------------------------------------------------------------------------
 * denoise.slang (:32-126) and PBRPathTracer::denoise_pass (src/PBRPathTracer.cpp:346-400): the
 * edge-avoiding a-trous wavelet over the pixel buffer, steps 1, 2, 4, 8, 16 (:26), ping-ponging between two
 * images; without denoising the image is radiance and alpha. px8: 8 words per pixel (radiance rgb, alpha,
 * depth, normal as two snorm16 octahedral coordinates, object id, albedo rgba8). out4: rgba per pixel.
 * ---------------------------------------------------------------------------------------------- */
static void dn_unpack_normal(uint32_t packed, float n[3])
{
    /* unpack_snorm_2x16 (utils/octahedral_map.slang:40-45), then the octahedral map (:25-35) */
    int16_t lo = (int16_t)(packed & 0xffffu), hi = (int16_t)(packed >> 16);
    float px = (float)lo / 32767.0f, py = (float)hi / 32767.0f;
    if(px < -1.0f) px = -1.0f;
    if(py < -1.0f) py = -1.0f;
    float x = px, y = py, z = 1.0f - fabsf(px) - fabsf(py);
    if(z < 0.0f)
    {
        float nx = (1.0f - fabsf(y)) * (x >= 0.0f ? 1.0f : -1.0f);
        float ny = (1.0f - fabsf(x)) * (y >= 0.0f ? 1.0f : -1.0f);
        x = nx;
        y = ny;
    }
    float inv = 1.0f / sqrtf((x * x + y * y) + z * z);
    n[0] = x * inv;
    n[1] = y * inv;
    n[2] = z * inv;
}

static void dn_albedo(uint32_t c, float a[3])
{
    a[0] = (float)(c & 0xffu) / 255.0f;
    a[1] = (float)((c >> 8) & 0xffu) / 255.0f;
    a[2] = (float)((c >> 16) & 0xffu) / 255.0f;
}

static void denoise_passes(const uint32_t *px8, uint32_t w, uint32_t h, int denoise, int passes, float *out4);

void orc_denoise(const uint32_t *px8, uint32_t w, uint32_t h, int denoise, float *out4)
{
    denoise_passes(px8, w, h, denoise, 5, out4);
}

void orc_denoise_passes(const uint32_t *px8, uint32_t w, uint32_t h, int passes, float *out4)
{
    denoise_passes(px8, w, h, 1, passes < 1 ? 1 : (passes > 5 ? 5 : passes), out4);
}

static void denoise_passes(const uint32_t *px8, uint32_t w, uint32_t h, int denoise, int passes, float *out4)
{
    const size_t n = (size_t)w * h;
    if(!denoise)
    {
        for(size_t i = 0; i < n; ++i) memcpy(out4 + 4 * i, px8 + 8 * i, 16);
        return;
    }
    static const float kern[25] = {1.0f / 256.0f, 1.0f / 64.0f, 3.0f / 128.0f, 1.0f / 64.0f, 1.0f / 256.0f,
                                   1.0f / 64.0f,  1.0f / 16.0f, 3.0f / 32.0f,  1.0f / 16.0f, 1.0f / 64.0f,
                                   3.0f / 128.0f, 3.0f / 32.0f, 9.0f / 64.0f,  3.0f / 32.0f, 3.0f / 128.0f,
                                   1.0f / 64.0f,  1.0f / 16.0f, 3.0f / 32.0f,  1.0f / 16.0f, 1.0f / 64.0f,
                                   1.0f / 256.0f, 1.0f / 64.0f, 3.0f / 128.0f, 1.0f / 64.0f, 1.0f / 256.0f};
    const float lum[3] = {0.298999994993f, 0.587000012398f, 0.114000000060f};
    const float color_phi = 1.0f / 0.9f, normal_phi = 1.0f / 0.3f, position_phi = 1.0f;
    const int steps[5] = {1, 2, 4, 8, 16};
    float *img[2];
    img[0] = (float *)malloc(n * 16);
    img[1] = (float *)malloc(n * 16);
    for(int pass = 0; pass < passes; ++pass)
    {
        const float *prev = pass ? img[(pass - 1) % 2] : NULL;
        float *dst = img[pass % 2];
        for(uint32_t y = 0; y < h; ++y)
            for(uint32_t x = 0; x < w; ++x)
            {
                const size_t c = (size_t)y * w + x;
                float cv[3], nc[3], ac[3], eff[3], cd[3];
                for(int k = 0; k < 3; ++k)
                {
                    float f;
                    memcpy(&f, px8 + 8 * c + k, 4);
                    cv[k] = prev ? prev[4 * c + k] : f;
                }
                float depth_c;
                memcpy(&depth_c, px8 + 8 * c + 4, 4);
                dn_unpack_normal(px8[8 * c + 5], nc);
                dn_albedo(px8[8 * c + 7], ac);
                for(int k = 0; k < 3; ++k)
                {
                    eff[k] = ac[k] > 0.01f ? ac[k] : 0.01f;
                    cd[k] = cv[k] / eff[k];
                }
                const float lum_c = (cd[0] * lum[0] + cd[1] * lum[1]) + cd[2] * lum[2];
                float sum[3] = {0.f, 0.f, 0.f}, cum = 0.f;
                for(int t = 0; t < 25; ++t)
                {
                    int ux = (int)x + (t % 5 - 2) * steps[pass], uy = (int)y + (t / 5 - 2) * steps[pass];
                    ux = ux < 0 ? 0 : (ux > (int)w - 1 ? (int)w - 1 : ux);
                    uy = uy < 0 ? 0 : (uy > (int)h - 1 ? (int)h - 1 : uy);
                    const size_t j = (size_t)uy * w + (size_t)ux;
                    float ct[3], at[3], nt[3], td[3], dj;
                    for(int k = 0; k < 3; ++k)
                    {
                        float f;
                        memcpy(&f, px8 + 8 * j + k, 4);
                        ct[k] = prev ? prev[4 * j + k] : f;
                    }
                    dn_albedo(px8[8 * j + 7], at);
                    for(int k = 0; k < 3; ++k) td[k] = ct[k] / (at[k] > 0.01f ? at[k] : 0.01f);
                    const float ld = lum_c - ((td[0] * lum[0] + td[1] * lum[1]) + td[2] * lum[2]);
                    float cw = expf(-(ld * ld) * color_phi);
                    cw = cw < 1.0f ? cw : 1.0f;
                    dn_unpack_normal(px8[8 * j + 5], nt);
                    const float dx = nc[0] - nt[0], dy = nc[1] - nt[1], dz = nc[2] - nt[2];
                    float nw = expf(-((dx * dx + dy * dy) + dz * dz) * normal_phi);
                    nw = nw < 1.0f ? nw : 1.0f;
                    memcpy(&dj, px8 + 8 * j + 4, 4);
                    const float pw = expf(-fabsf(depth_c - dj) * position_phi);
                    const float wt = ((cw * nw) * pw) * kern[t];
                    for(int k = 0; k < 3; ++k) sum[k] = sum[k] + td[k] * wt;
                    cum = cum + wt;
                }
                for(int k = 0; k < 3; ++k) dst[4 * c + k] = (sum[k] / cum) * eff[k];
                dst[4 * c + 3] = 0.f;
            }
    }
    memcpy(out4, img[(passes - 1) % 2], n * 16);/* pass k writes image k % 2 */
    free(img[0]);
    free(img[1]);
}

/* The value of x as IEEE binary16 stores it: round to nearest, ties to even, overflow to infinity. The
 * quotient by the spacing of halves at x's binade is exact in double, so nearbyint (default rounding mode)
 * does the rounding; subnormal halves have spacing 2^-24. */
float orc_half_value(float x)
{
    if(!(fabsf(x) < INFINITY) || x == 0.f) return x;
    int k;
    frexp(x, &k);
    const int q = (k - 1 >= -14) ? (k - 1) - 10 : -24;
    const double r = ldexp(nearbyint(ldexp((double)x, -q)), q);
    if(fabs(r) > 65504.0) return copysignf(INFINITY, x);
    return (float)r;
}

/* bloom_composition.slang:34-45 with neither bloom nor motion blur (PBRPathTracer.cpp:405-465):
 * utils/tonemap.slang tonemap_exposure, pow(., 1 / gamma) with the single-precision quotient, alpha passed on,
 * every channel stored in the R16G16B16A16_SFLOAT target (PBRPathTracer.cpp:819) */
void orc_tonemap(const float *in4, size_t n, float exposure, float gamma, float *out4)
{
    const float inv_gamma = 1.0f / gamma;
    for(size_t i = 0; i < n; ++i)
    {
        for(int k = 0; k < 3; ++k)
        {
            float r = 1.0f - expf(-in4[4 * i + k] * exposure);
            if(inv_gamma != 1.0f) r = powf(r, inv_gamma);
            out4[4 * i + k] = orc_half_value(r);
        }
        out4[4 * i + 3] = orc_half_value(in4[4 * i + 3]);
    }
}
