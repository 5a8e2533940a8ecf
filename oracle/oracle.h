/*
 * CPU oracle for the Raymond rendering path — TEST INFRASTRUCTURE ONLY.
 *
 * A plain C++ restatement of the reference's WGSL trace and denoise shaders
 * (src/shaders.mjs) and of its CPU raycast (src/Scene.mjs:202-283), used by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg to check
 * and time the HIP kernels. Nothing in the product links, loads or calls it.
 *
 * Arithmetic is FP32 with the shader's operation order, compiled without
 * FMA contraction, so that hit records are the bit-exact contract for the
 * GPU traversal; oracle_trace64 is the same traversal in FP64 (the
 * reference's JS raycast semantics) for error bars.
 */
#pragma once
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A packed scene as the reference's Renderer uploads it (all host memory). */
typedef struct OracleScene {
    const uint32_t* prims;       /* 16 u32 per primitive slot (64 B) */
    uint32_t prim_slots;
    const uint32_t* nodes;       /* 8 u32 per node (32 B) */
    uint32_t node_count;
    const uint32_t* materials;   /* 24 u32 per material (96 B) */
    uint32_t material_count;
    const float* attributes;     /* vec4s */
    uint32_t attribute_vec4s;
    const uint32_t* lights;      /* light prims, volume prims, cdf bits */
    uint32_t lights_len;
    const uint16_t* env_rgba;    /* half floats, 4 per texel */
    uint32_t env_width, env_height;
    const float* env_table;      /* table_width × (table_height + 1) */
    uint32_t table_width, table_height;
    const uint8_t* textures;     /* sRGB RGBA8 layers, tex_width × tex_height each */
    uint32_t tex_width, tex_height, tex_layers;
    const uint8_t* normal_maps;  /* RGBA8 layers */
    uint32_t nmap_width, nmap_height, nmap_layers;
} OracleScene;

/* Rays: 8 floats each (ox, oy, oz, t_max, dx, dy, dz, unused).
 * Hits: 4 words each (t f32, prim u32, u f32, v f32).
 * counters (optional): 2 u32 per ray: node records tested, primitive records tested. */
void oracle_trace(const OracleScene* s, const float* rays, uint32_t n, int any_hit,
                  uint32_t* hits, uint32_t* counters, int threads);
void oracle_trace64(const OracleScene* s, const float* rays, uint32_t n, double* hits, int threads);

/* Exhaustive closest hit over primitives [0, prim_count), each tested with
 * an unbounded t_max: 2 words per ray (t, primitive or 0xffffffff). */
void oracle_brute_force(const OracleScene* s, const float* rays, uint32_t n, uint32_t prim_count, uint32_t* hits);

/* Pixel-centre pinhole rays for a parameter block (camera_ray, src/shaders.mjs:1148-1166). */
void oracle_camera_rays(const uint8_t* params, float* rays);

/* One frame of `main` (src/shaders.mjs:1838-1895) over rows [y0, y1). */
void oracle_main(const OracleScene* s, const uint8_t* params, float* accum, float* guide_albedo,
                 float* guide_normal, uint32_t y0, uint32_t y1, int threads);

/* Segment statistics of main's camera paths over a pixel subsample (every
 * `stride`-th pixel): counts[0] closest-hit rays, [1] shadow rays, [2]
 * node records tested, [3] primitive records tested, [4] surface hits,
 * [5] smooth-triangle hits, [6] samples. */
void oracle_path_stats(const OracleScene* s, const uint8_t* params, uint32_t stride, uint64_t* counts);

/* Light tracing (src/shaders.mjs:1630-1814) for photons [first, first+count):
 * splats u32×4 per pixel, photon_heads (2^20 + 1) u32, photons 12 f32 each. */
void oracle_light_trace(const OracleScene* s, const uint8_t* params, const float* emitters,
                        uint32_t first, uint32_t count, uint32_t* splats, uint32_t* photon_heads,
                        float* photons);
/* The same with a light-path budget: `paths` light paths a frame (the
 * normalisation of each photon's power) and room for `capacity` photons. */
void oracle_light_trace_budget(const OracleScene* s, const uint8_t* params, const float* emitters,
                               uint32_t first, uint32_t count, uint32_t paths, uint32_t capacity,
                               uint32_t* splats, uint32_t* photon_heads, float* photons);
void oracle_resolve_splats(const uint8_t* params, uint32_t* splats, float* accum);
void oracle_photon_gather(const OracleScene* s, const uint8_t* params, const uint32_t* photon_heads,
                          const float* photons, float* accum, int threads);

/* The denoiser (src/shaders.mjs:1916-2098): prepare, atrous at step, compose. */
void oracle_prepare(const uint8_t* params, const float* accum, const float* guide_albedo,
                    const float* guide_normal, float* light_out);
void oracle_atrous(const uint8_t* params, uint32_t step, const float* guide_normal,
                   const float* light_in, float* light_out);
void oracle_compose(const uint8_t* params, const float* guide_albedo, const float* light_in,
                    float* out);

/* The whole denoiser (prepare, à-trous 1..16, compose) in FP64 from the FP32
 * sums: the error reference for the GPU's FP32 and FP16-guide passes. */
void oracle_denoise64(const uint8_t* params, const float* accum, const float* guide_albedo,
                      const float* guide_normal, double* out);

/* DISPLAY_SHADER (src/shaders.mjs:2100-2152) as RGBA8 per pixel. */
void oracle_display(const uint8_t* params, const float* accum, const float* denoised, uint8_t* out);

/* Building blocks, exposed for known-answer and property tests. */
uint32_t oracle_pcg(uint32_t v);
uint32_t oracle_owen(uint32_t x, uint32_t seed);
uint32_t oracle_sobol2(uint32_t i);
uint32_t oracle_reverse_bits(uint32_t x);
/* The first n rand() draws of main's sample k at pixel index, dims from dim0. */
void oracle_qmc_draws(uint32_t pixel, uint32_t sample, uint32_t dim0, uint32_t n, int pairs, float* out);
/* intersect() of one primitive: out = (t or -1, u, v). */
void oracle_intersect(const OracleScene* s, uint32_t prim, const float* ray, float t_max, float* out);
/* eval_bsdf / sample_bsdf for a material record at a frame where n = (0,0,1). */
void oracle_eval_bsdf(const uint32_t* material, const float* base, const float* n, const float* wo,
                      const float* wi, float* out4);
float oracle_fresnel_dielectric(float cos_i, float eta);
/* GGX for a roughness (alpha = max(r², 0.002)): D(n·h), G1(n·v), G1(n·l). */
void oracle_ggx_terms(float roughness, float nh, float nv, float nl, float* out3);
/* sample_ggx's half vector about n = (0, 0, 1) for the uniforms (u1, u2). */
void oracle_ggx_sample(float roughness, float u1, float u2, float* h3);
/* sample_bsdf on a front surface with n = ng, drawing from a PCG stream
 * started at seed: wi[3], weight[3], pdf, delta (0/1), ok (0/1). */
void oracle_sample_bsdf(const uint32_t* material, const float* base, const float* n, const float* wo,
                        uint32_t seed, float* out9);
void oracle_wavelength_rgb(float l, float* rgb);
float oracle_env_pdf(const OracleScene* s, const uint8_t* params, const float* d);
void oracle_env_radiance(const OracleScene* s, const uint8_t* params, const float* d, float* rgb);

#ifdef __cplusplus
}
#endif
