// Raymond's frame kernels for gfx950 (one wave64 per 8×8 pixel tile). This
// file holds the raycast, the denoiser and the display pass:
//   k_raycast        closest-hit / any-hit hit records for a batch of rays
//   k_guides / k_prepare / k_atrous_staged / k_atrous / k_compose
//                    the à-trous denoiser
//   k_finish         resolve, guides and prepare in one pass once 4 or more
//                    samples have accumulated
//   k_display        the display pass's RGBA8 image
// The path-tracing passes are in main.hpp (k_main), batch.hip (k_main_batch)
// and schedules.hip (compacted, persistent, wavefront); light tracing, the
// photon gather and the resolve in light.hip; the path integrator they share
// in path.hpp. Each kernel restates the reference entry point cited beside
// it (src/shaders.mjs). Traversal stacks live in LDS with a global spill.
#include <hip/hip_fp16.h>

#include "path.hpp"

namespace rm {

// ------------------------------------------------------------------ raycast
// trace() over a ray batch: rays are 8 floats (o, t_max, d, pad); hits are
// (t, prim, u, v) as 4 words.
template <bool ANY, int TREE>
__global__ __launch_bounds__(BLOCK_THREADS) void k_raycast(DevScene S, const float4* __restrict__ rays, uint32_t n,
                                                           uint4* __restrict__ hits, uint32_t* spill, uint32_t slots,
                                                           uint2* counters) {
    __shared__ uint2 lds[TRACE_LDS_DEPTH * BLOCK_THREADS];
    uint32_t i = blockIdx.x * BLOCK_THREADS + threadIdx.x;
    if (i >= n) return;
    auto stack = BlockStack<TRACE_LDS_DEPTH, TREE>::make(lds, spill, i, slots);
    float4 a = rays[(size_t)i * 2], b = rays[(size_t)i * 2 + 1];
    Hit h;
    if (counters) {
        Count c;
        h = trace<ANY>(S, xyz(a), xyz(b), a.w, stack, c);
        counters[i] = make_uint2(c.n, c.p);
    } else {
        NoCount c;
        h = trace<ANY>(S, xyz(a), xyz(b), a.w, stack, c);
    }
    hits[i] = make_uint4(__float_as_uint(h.t), h.prim, __float_as_uint(h.u), __float_as_uint(h.v));
}

// ------------------------------------------------------------------ denoiser
// src/shaders.mjs:1916-2098
constexpr float KERNEL0 = 0.375f, KERNEL1 = 0.25f, KERNEL2 = 0.0625f;
RM_D float kernel_w(int k) { return k == 0 ? KERNEL0 : (k == 1 ? KERNEL1 : KERNEL2); }

RM_D V3 albedo_at(const float4* galb, uint32_t i, float n) {
    float4 g = galb[i];
    return vmax(v3(g.x, g.y, g.z) / n, s3(0.01f));
}
// The average normal (unit, or zero where none) and average distance.
RM_D float4 normal_at(const float4* gnorm, uint32_t i, float n) {
    float4 g = gnorm[i];
    V3 v = {g.x, g.y, g.z};
    V3 nn = dot(v, v) > 0.0f ? normalize(v) : v3(0, 0, 0);
    return make_float4(nn.x, nn.y, nn.z, g.w / n);
}
RM_D float similarity(float4 p, float4 q, float step) {
    if (p.w <= 0.0f || q.w <= 0.0f) return (p.w <= 0.0f && q.w <= 0.0f) ? 1.0f : 0.0f;
    if (p.x * p.x + p.y * p.y + p.z * p.z == 0.0f) return 0.0f;
    float wn = powf(fmaxf(dot(v3(p.x, p.y, p.z), v3(q.x, q.y, q.z)), 0.0f), 64.0f);
    float wz = expf(-fabsf(p.w - q.w) / (p.w * 0.02f * step + 1e-6f));
    return wn * wz;
}
RM_D float pow64(float x) {
    x *= x; x *= x; x *= x; x *= x; x *= x; x *= x;
    return x;
}

// How the per-frame guide (unit normal, mean distance) is stored for the
// 25-tap passes: FP32, or FP16 (half the bytes every tap reads). The sums
// it comes from stay FP32, and the passes compute in FP32 either way.
struct GuidesF32 {
    using T = float4;
    RM_D static float4 get(const T* g, uint32_t i) { return g[i]; }
    RM_D static T put(float4 v) { return v; }
};
struct GuidesF16 {
    using T = ushort4;
    RM_D static float wide(unsigned short h) { return __half2float(__ushort_as_half(h)); }
    // Round to nearest; distances past FP16's range clamp to its largest finite value.
    RM_D static unsigned short narrow(float f) {
        return __half_as_ushort(__float2half_rn(clampf(f, -65504.0f, 65504.0f)));
    }
    RM_D static float4 get(const T* g, uint32_t i) {
        const ushort4 u = g[i];
        return make_float4(wide(u.x), wide(u.y), wide(u.z), wide(u.w));
    }
    RM_D static T put(float4 v) { return make_ushort4(narrow(v.x), narrow(v.y), narrow(v.z), narrow(v.w)); }
};

// Each pixel's normal_at() once per frame, so the 25-tap passes read it
// rather than renormalising every neighbour's guide.
template <class G>
__global__ __launch_bounds__(BLOCK_THREADS) void k_guides(Params P, const float4* __restrict__ gnorm,
                                                          typename G::T* __restrict__ guides) {
    uint32_t i = blockIdx.x * BLOCK_THREADS + threadIdx.x;
    if (i >= P.width * P.height) return;
    guides[i] = G::put(normal_at(gnorm, i, (float)max(P.total_samples, 1u)));
}

// Prepare's variance once 4 or more samples have accumulated: from the
// pixel's own luminance moments, demodulated by its albedo.
RM_D float temporal_variance(float4 ac, V3 own_albedo, float n) {
    float a = luminance(own_albedo);
    float mean = luminance(v3(ac.x, ac.y, ac.z)) / n;
    float spread = fmaxf(ac.w / n - mean * mean, 0.0f) * n / (n - 1.0f);
    return spread / n / (a * a);
}

// src/shaders.mjs:1974-2021. The pixel's albedo (albedo_at) also goes to
// albedo_out for k_compose, so that the passes after this one no longer read
// the accumulation buffers the next frame writes.
template <class G>
__global__ __launch_bounds__(BLOCK_THREADS) void k_prepare(Params P, const float4* __restrict__ accum,
                                                           const float4* __restrict__ galb,
                                                           const typename G::T* __restrict__ guides, float4* light_out,
                                                           float4* __restrict__ albedo_out) {
    uint32_t x, y;
    if (!stencil_pixel(P.width, P.height, x, y)) return;
    const uint32_t i = y * P.width + x;
    const float n = (float)max(P.total_samples, 1u);
    float4 ac = accum[i];
    const V3 own_albedo = albedo_at(galb, i, n);
    albedo_out[i] = make_float4(own_albedo.x, own_albedo.y, own_albedo.z, 0.0f);
    V3 light = (v3(ac.x, ac.y, ac.z) / n) / own_albedo;
    float4 gp = G::get(guides, i);
    float variance;
    if (n >= 4.0f) {
        variance = temporal_variance(ac, own_albedo, n);
    } else {
        float sum = 0.0f, sum2 = 0.0f, weight = 0.0f;
        for (int dy = -2; dy <= 2; dy++)
            for (int dx = -2; dx <= 2; dx++) {
                int xx = (int)x + dx, yy = (int)y + dy;
                if (xx < 0 || yy < 0 || xx >= (int)P.width || yy >= (int)P.height) continue;
                uint32_t j = (uint32_t)yy * P.width + (uint32_t)xx;
                float w = similarity(gp, G::get(guides, j), 1.0f);
                float4 aj = accum[j];
                float l = luminance((v3(aj.x, aj.y, aj.z) / n) / albedo_at(galb, j, n));
                sum += w * l;
                sum2 += w * l * l;
                weight += w;
            }
        float mean = sum / fmaxf(weight, 1e-6f);
        variance = fmaxf(sum2 / fmaxf(weight, 1e-6f) - mean * mean, 0.0f) + 0.01f * mean * mean;
    }
    light_out[i] = make_float4(light.x, light.y, light.z, variance);
}

// The path stream's work between two path passes once 4 or more samples
// have accumulated, when prepare reads only its own pixel: the splat and
// gather resolve (RESOLVE, with light tracing), the pixel's guide and
// prepare, in one pass over the pixels. The same arithmetic as k_resolve,
// k_guides and k_prepare, with each buffer read once and two launches fewer.
template <class G, bool RESOLVE>
__global__ __launch_bounds__(BLOCK_THREADS) void k_finish(Params P, uint4* splats, const float4* __restrict__ gathered,
                                                          float4* __restrict__ accum, const float4* __restrict__ galb,
                                                          const float4* __restrict__ gnorm,
                                                          typename G::T* __restrict__ guides,
                                                          float4* __restrict__ light_out,
                                                          float4* __restrict__ albedo_out) {
    const uint32_t i = blockIdx.x * BLOCK_THREADS + threadIdx.x;
    if (i >= P.width * P.height) return;
    const float n = (float)max(P.total_samples, 1u);
    float4 ac = accum[i];
    if (RESOLVE) {
        ac = resolve_add(P, splats, gathered, nullptr, i, ac);
        accum[i] = ac;
    }
    guides[i] = G::put(normal_at(gnorm, i, n));
    const V3 own_albedo = albedo_at(galb, i, n);
    albedo_out[i] = make_float4(own_albedo.x, own_albedo.y, own_albedo.z, 0.0f);
    const V3 light = (v3(ac.x, ac.y, ac.z) / n) / own_albedo;
    light_out[i] = make_float4(light.x, light.y, light.z, temporal_variance(ac, own_albedo, n));
}

// src/shaders.mjs:2023-2086: one pixel of an à-trous pass at `step`. The 25
// taps' weights are the reference's B3 × max(n·n', 0)^64 ×
// exp(-|Δz| / (0.02 z step)) × exp(-|Δl| / σ), with the power by six
// squarings, the two exponentials merged into one exp2, and the centre's
// divisions hoisted to reciprocals: a few ulps from the reference's per-tap
// pow/exp/divide, several times fewer instructions. σ comes from the 3×3
// (step 1) neighbourhood's variance. tap(dx, dy) gives the light and guide
// of the pixel (dx, dy) taps away, and var(dx, dy) the variance of the pixel
// (dx, dy) away; both are asked only for pixels inside the image.
//
// Unlike the path pass, the denoiser is held to a relative tolerance, not
// bit for bit, so its multiply-adds are fused and each weight's exponential
// is the hardware exp2 (weights below 2^-126 become 0): with the taps staged
// in LDS the passes are bound by arithmetic, and this makes them 8% faster
// and the frame 1.2%.
template <class TAP, class VAR>
RM_D float4 atrous_pixel(int x, int y, int W, int H, int step, float4 centre, float4 gp, TAP&& tap, VAR&& var) {
#pragma clang fp contract(fast)
    const float lp = luminance(v3(centre.x, centre.y, centre.z));
    float variance = 0.0f, variance_weight = 0.0f;
    for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
            int xx = x + dx, yy = y + dy;
            if (xx >= 0 && yy >= 0 && xx < W && yy < H) {
                float w = kernel_w(abs(dx)) * kernel_w(abs(dy));
                variance += w * var(dx, dy);
                variance_weight += w;
            }
        }
    const float LOG2E = 1.4426950408889634f;
    const float sigma = 4.0f * sqrtf(variance / variance_weight) + 1e-4f;
    const float k_l = LOG2E / sigma;
    const bool sky = gp.w <= 0.0f;
    const bool no_normal = gp.x * gp.x + gp.y * gp.y + gp.z * gp.z == 0.0f;
    const float k_z = LOG2E / (gp.w * 0.02f * (float)step + 1e-6f);
    V3 sum = {0, 0, 0};
    float sum_variance = 0.0f, weight = 0.0f;
#pragma unroll
    for (int dy = -2; dy <= 2; dy++) {
        int yy = y + dy * step;
        if (yy < 0 || yy >= H) continue;
#pragma unroll
        for (int dx = -2; dx <= 2; dx++) {
            int xx = x + dx * step;
            if (xx < 0 || xx >= W) continue;
            float4 q, g;
            tap(dx, dy, q, g);
            float e = -fabsf(lp - luminance(v3(q.x, q.y, q.z))) * k_l;
            float sim;
            if (sky || g.w <= 0.0f) {
                sim = (sky && g.w <= 0.0f) ? 1.0f : 0.0f;
            } else if (no_normal) {
                sim = 0.0f;
            } else {
                sim = pow64(fmaxf(gp.x * g.x + gp.y * g.y + gp.z * g.z, 0.0f));
                e -= fabsf(gp.w - g.w) * k_z;
            }
            float w = kernel_w(abs(dx)) * kernel_w(abs(dy)) * sim * __builtin_amdgcn_exp2f(e);
            sum = sum + w * v3(q.x, q.y, q.z);
            sum_variance += w * w * q.w;
            weight += w;
        }
    }
    if (!(weight > 0.0f)) return centre;
    V3 r = sum / weight;
    return make_float4(r.x, r.y, r.z, sum_variance / (weight * weight));
}

// An à-trous pass reading its taps straight from the buffers, for steps 8
// and 16: a wave's tap is one coalesced load, and the neighbouring
// workgroups of an XCD read the same lines.
template <class G>
RM_D void atrous_direct(const Params& P, int step, const typename G::T* __restrict__ guides,
                        const float4* __restrict__ light_in, float4* __restrict__ light_out, uint32_t ux,
                        uint32_t uy) {
    const int x = (int)ux, y = (int)uy;
    const uint32_t W = P.width;
    const uint32_t i = uy * W + ux;
    auto tap = [&](int dx, int dy, float4& q, float4& g) {
        const uint32_t j = (uint32_t)(y + dy * step) * W + (uint32_t)(x + dx * step);
        q = light_in[j];
        g = G::get(guides, j);
    };
    auto var = [&](int dx, int dy) { return light_in[(uint32_t)(y + dy) * W + (uint32_t)(x + dx)].w; };
    light_out[i] = atrous_pixel(x, y, (int)P.width, (int)P.height, step, light_in[i], G::get(guides, i), tap, var);
}
template <class G>
__global__ __launch_bounds__(BLOCK_THREADS) void k_atrous(Params P, int step, const typename G::T* __restrict__ guides,
                                                          const float4* __restrict__ light_in,
                                                          float4* __restrict__ light_out) {
    uint32_t ux, uy;
    if (stencil_pixel(P.width, P.height, ux, uy)) atrous_direct<G>(P, step, guides, light_in, light_out, ux, uy);
}

// Several jobs' à-trous pass at `step` in one launch (rm_frame_batch): the
// grid is each job's stencil grid, one job after another, and each block
// runs k_atrous's block of the same index for its job, reading its taps
// directly rather than from LDS. STEP is the step as a compile-time
// constant (0: the `step` argument), as in the single-job kernels: the
// staged ones (steps 1, 2 and 4) know their step at compile time and
// k_atrous (8 and 16) doesn't, and the multiply-adds the compiler fuses in
// the weights differ between the two, so matching them keeps every pixel
// the same bit for bit. A job reads its a and writes its b when `parity` is
// 0, the other way round when 1. The job table is the kernel argument, read
// by scalar loads.
RM_D uint32_t denoise_job(const DenoiseTable& T) {
    uint32_t j = 0;
    while (j + 1u < T.n && blockIdx.x >= T.begin[j + 1u]) ++j;
    return j;
}
template <class G, int STEP>
__global__ __launch_bounds__(BLOCK_THREADS) void k_atrous_batch(DenoiseTable T, int step, int parity) {
    const uint32_t j = denoise_job(T);
    const DenoiseJob& J = T.job[j];
    uint32_t ux, uy;
    if (!stencil_pixel_at(J.P.width, J.P.height, blockIdx.x - T.begin[j], ux, uy)) return;
    atrous_direct<G>(J.P, STEP > 0 ? STEP : step, static_cast<const typename G::T*>(J.guides), parity ? J.b : J.a,
                     parity ? J.a : J.b, ux, uy);
}

// The staged à-trous pass, for steps 1, 2 and 4. At step s the taps of
// pixel (x, y) are the pixels x + s·i, y + s·j (|i|, |j| <= 2), all in the
// same phase (x mod s, y mod s), where the pass is a dense 5×5 filter. A
// workgroup takes PG = s adjacent x-phases of one y-phase, a TU×TV tile of
// each, stages them with their 2-pixel apron in LDS once and reads every tap
// from there: 1.6–2.3 staged loads of each buffer per pixel rather than 25.
// A staged row holds PG neighbouring pixels, so its loads stay coalesced.
// Consecutive logical blocks are the phase groups of one tile, and each XCD
// (workgroups go to the 8 XCDs round-robin) takes one contiguous run of them.
// At 1920×1080, alone, steps 1, 2 and 4 take 60, 67 and 68 µs against 106,
// 101 and 90 µs for the direct pass. Staged with eight phases a workgroup,
// steps 8 and 16 take 69 and 72 µs against 80 and 81 alone, but need 40 VGPRs
// and 24 KB of LDS, and beside the next frame's path pass they made the frame
// 1% slower, so those steps stay direct.
template <int PG> struct AtrousShape;
template <> struct AtrousShape<1> { static constexpr int TU = 16, TV = 16; };
template <> struct AtrousShape<2> { static constexpr int TU = 8, TV = 16; };
template <> struct AtrousShape<4> { static constexpr int TU = 8, TV = 8; };

struct AtrousGrid {
    uint32_t groups;    // phase groups: s / PG x-groups × s y-phases
    uint32_t tiles_u;   // tiles across one phase
    uint32_t tiles;     // tiles per phase
    uint32_t blocks;    // logical blocks: groups × tiles
    uint32_t launched;  // blocks rounded up to a multiple of the XCD count
};
template <int PG>
inline AtrousGrid atrous_grid(uint32_t width, uint32_t height, uint32_t step) {
    using A = AtrousShape<PG>;
    const uint32_t u = (width + step - 1u) / step, v = (height + step - 1u) / step;
    AtrousGrid g;
    g.groups = step / PG * step;
    g.tiles_u = (u + A::TU - 1u) / A::TU;
    g.tiles = g.tiles_u * ((v + A::TV - 1u) / A::TV);
    g.blocks = g.groups * g.tiles;
    g.launched = (g.blocks + XCDS - 1u) / XCDS * XCDS;
    return g;
}

// PG is the step itself (1, 2 or 4): a workgroup takes all s x-phases of one
// y-phase, so the step is a compile-time constant of each instance.
template <class G, int PG>
__global__ __launch_bounds__(BLOCK_THREADS) void k_atrous_staged(Params P, AtrousGrid grid,
                                                                 const typename G::T* __restrict__ guides,
                                                                 const float4* __restrict__ light_in,
                                                                 float4* __restrict__ light_out) {
    constexpr int TU = AtrousShape<PG>::TU, TV = AtrousShape<PG>::TV;
    static_assert(PG * TU * TV == BLOCK_THREADS, "a thread per staged output");
    constexpr int SU = TU + 4, SV = TV + 4, N = PG * SU * SV;
    __shared__ float4 tile_light[N];
    __shared__ float4 tile_guide[N];
    const uint32_t logical = (blockIdx.x % XCDS) * (grid.launched / XCDS) + blockIdx.x / XCDS;
    if (logical >= grid.blocks) return;
    constexpr int s = PG;
    const uint32_t group = logical % grid.groups, t = logical / grid.groups;
    const int xgroups = s / PG;
    const int gx = (int)(group % (uint32_t)xgroups) * PG, py = (int)(group / (uint32_t)xgroups);
    const int u0 = (int)(t % grid.tiles_u) * TU, v0 = (int)(t / grid.tiles_u) * TV;
    const int W = (int)P.width, H = (int)P.height;
    for (int e = (int)threadIdx.x; e < N; e += BLOCK_THREADS) {
        const int a = e % PG, cu = (e / PG) % SU, rv = e / (PG * SU);
        const int xx = gx + a + s * (u0 + cu - 2), yy = py + s * (v0 + rv - 2);
        if (xx >= 0 && yy >= 0 && xx < W && yy < H) {
            const uint32_t j = (uint32_t)yy * P.width + (uint32_t)xx;
            tile_light[e] = light_in[j];
            tile_guide[e] = G::get(guides, j);
        }
    }
    __syncthreads();
    const int th = (int)threadIdx.x;
    const int a = th % PG, tu = (th / PG) % TU, tv = th / (PG * TU);
    const int x = gx + a + s * (u0 + tu), y = py + s * (v0 + tv);
    if (x >= W || y >= H) return;
    const int c = a + PG * ((tu + 2) + SU * (tv + 2));
    auto tap = [&](int dx, int dy, float4& q, float4& g) {
        q = tile_light[c + PG * (dx + SU * dy)];
        g = tile_guide[c + PG * (dx + SU * dy)];
    };
    // The variance neighbourhood is in the tile only at step 1 (decided at
    // compile time: a run-time choice between the tile and the buffer became
    // one flat load, which waits on both memory counters).
    auto var = [&](int dx, int dy) {
        if constexpr (s == 1) return tile_light[c + dx + SU * dy].w;
        else return light_in[(uint32_t)(y + dy) * P.width + (uint32_t)(x + dx)].w;
    };
    light_out[(uint32_t)y * P.width + (uint32_t)x] = atrous_pixel(x, y, W, H, s, tile_light[c], tile_guide[c], tap, var);
}

// src/shaders.mjs:2088-2098, with the albedo k_prepare recorded.
RM_D void compose_pixel(const Params& P, const float4* __restrict__ albedo, const float4* __restrict__ light_in,
                        float4* out, uint32_t x, uint32_t y) {
    const uint32_t i = y * P.width + x;
    float4 l = light_in[i], a = albedo[i];
    V3 r = v3(l.x, l.y, l.z) * v3(a.x, a.y, a.z);
    out[i] = make_float4(r.x, r.y, r.z, 0.0f);
}
__global__ __launch_bounds__(BLOCK_THREADS) void k_compose(Params P, const float4* __restrict__ albedo,
                                                           const float4* __restrict__ light_in, float4* out) {
    uint32_t x, y;
    tile_pixel(x, y);
    if (x < P.width && y < P.height) compose_pixel(P, albedo, light_in, out, x, y);
}
// The jobs' compose in one launch, over the à-trous batch's grid; each job's
// light is in its b (the fifth pass's output).
__global__ __launch_bounds__(BLOCK_THREADS) void k_compose_batch(DenoiseTable T) {
    const uint32_t j = denoise_job(T);
    const DenoiseJob& J = T.job[j];
    uint32_t x, y;
    if (stencil_pixel_at(J.P.width, J.P.height, blockIdx.x - T.begin[j], x, y))
        compose_pixel(J.P, J.albedo, J.b, J.out, x, y);
}

// DISPLAY_SHADER's fragment (src/shaders.mjs:2100-2152) for every pixel at
// once: what render() draws to the canvas, as RGBA8 (each channel
// round(clamp(x, 0, 1)·255), as an 8-bit unorm target stores it).
RM_D float aces1(float x) {
    return clampf((x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f), 0.0f, 1.0f);
}
RM_D float srgb1(float c) {
    const float x = clampf(c, 0.0f, 1.0f);
    return x <= 0.0031308f ? x * 12.92f : 1.055f * powf(x, 1.0f / 2.4f) - 0.055f;
}
RM_D uint32_t unorm8(float x) { return (uint32_t)rintf(clampf(x, 0.0f, 1.0f) * 255.0f); }

__global__ __launch_bounds__(BLOCK_THREADS) void k_display(Params P, const float4* __restrict__ accum,
                                                           const float4* __restrict__ denoised,
                                                           uchar4* __restrict__ out) {
    const uint32_t i = blockIdx.x * BLOCK_THREADS + threadIdx.x;
    if (i >= P.width * P.height) return;
    V3 rgb = {0.0f, 0.0f, 0.0f};
    if (P.total_samples != 0u) {
        const float4 a = accum[i];
        const V3 c = v3(a.x, a.y, a.z) / (float)P.total_samples;
        if (P.view == VIEW_COLOR) {
            const float4 d = denoised[i];
            const V3 e = (P.denoise != 0u ? v3(d.x, d.y, d.z) : c) * P.exposure;
            rgb = {srgb1(aces1(e.x)), srgb1(aces1(e.y)), srgb1(aces1(e.z))};
        } else if (P.view == VIEW_ALBEDO) {
            rgb = {srgb1(c.x), srgb1(c.y), srgb1(c.z)};
        } else if (P.view == VIEW_NORMAL) {
            rgb = {c.x * 0.5f + 0.5f, c.y * 0.5f + 0.5f, c.z * 0.5f + 0.5f};
        } else {
            const float g = c.x <= 0.0f ? 0.0f : 1.0f / (1.0f + c.x / P.focus_distance);
            rgb = {g, g, g};
        }
    }
    out[i] = make_uchar4((unsigned char)unorm8(rgb.x), (unsigned char)unorm8(rgb.y), (unsigned char)unorm8(rgb.z), 255);
}

__global__ void k_fill_u32(uint32_t* p, uint32_t value, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = value;
}

}  // namespace rm

// ------------------------------------------------------------------ launchers
namespace rm {

void launch_raycast(const DevScene& S, const float4* rays, uint32_t n, uint4* hits, bool any_hit, int tree,
                    uint32_t* spill, uint32_t slots, uint2* counters, hipStream_t stream) {
    dim3 grid((n + BLOCK_THREADS - 1) / BLOCK_THREADS);
    with_tree(tree, [&](auto t) {
        constexpr int TREE = decltype(t)::value;
        if (any_hit)
            hipLaunchKernelGGL((k_raycast<true, TREE>), grid, dim3(BLOCK_THREADS), 0, stream, S, rays, n, hits, spill,
                               slots, counters);
        else
            hipLaunchKernelGGL((k_raycast<false, TREE>), grid, dim3(BLOCK_THREADS), 0, stream, S, rays, n, hits, spill,
                               slots, counters);
    });
}

void launch_guides(const Params& P, const float4* gnorm, void* guides, bool fp16, hipStream_t stream) {
    uint32_t n = P.width * P.height;
    dim3 grid((n + BLOCK_THREADS - 1) / BLOCK_THREADS);
    if (fp16)
        hipLaunchKernelGGL(k_guides<GuidesF16>, grid, dim3(BLOCK_THREADS), 0, stream, P, gnorm,
                           static_cast<ushort4*>(guides));
    else
        hipLaunchKernelGGL(k_guides<GuidesF32>, grid, dim3(BLOCK_THREADS), 0, stream, P, gnorm,
                           static_cast<float4*>(guides));
}

template <class G>
static void launch_finish_g(const Params& P, uint4* splats, const float4* gathered, float4* accum,
                            const float4* galb, const float4* gnorm, void* guides, float4* out, float4* albedo_out,
                            dim3 grid, hipStream_t stream) {
    auto* g = static_cast<typename G::T*>(guides);
    if (splats)
        hipLaunchKernelGGL((k_finish<G, true>), grid, dim3(BLOCK_THREADS), 0, stream, P, splats, gathered, accum,
                           galb, gnorm, g, out, albedo_out);
    else
        hipLaunchKernelGGL((k_finish<G, false>), grid, dim3(BLOCK_THREADS), 0, stream, P, splats, gathered, accum,
                           galb, gnorm, g, out, albedo_out);
}

void launch_finish(const Params& P, uint4* splats, const float4* gathered, float4* accum, const float4* galb,
                   const float4* gnorm, void* guides, bool fp16, float4* out, float4* albedo_out,
                   hipStream_t stream) {
    uint32_t n = P.width * P.height;
    dim3 grid((n + BLOCK_THREADS - 1) / BLOCK_THREADS);
    if (fp16)
        launch_finish_g<GuidesF16>(P, splats, gathered, accum, galb, gnorm, guides, out, albedo_out, grid, stream);
    else
        launch_finish_g<GuidesF32>(P, splats, gathered, accum, galb, gnorm, guides, out, albedo_out, grid, stream);
}

void launch_prepare(const Params& P, const float4* accum, const float4* galb, const void* guides, bool fp16,
                    float4* out, float4* albedo_out, hipStream_t stream) {
    dim3 grid(stencil_blocks(P.width, P.height));
    if (fp16)
        hipLaunchKernelGGL(k_prepare<GuidesF16>, grid, dim3(BLOCK_THREADS), 0, stream, P, accum, galb,
                           static_cast<const ushort4*>(guides), out, albedo_out);
    else
        hipLaunchKernelGGL(k_prepare<GuidesF32>, grid, dim3(BLOCK_THREADS), 0, stream, P, accum, galb,
                           static_cast<const float4*>(guides), out, albedo_out);
}

template <int PG>
static void launch_atrous_pg(const Params& P, const void* guides, bool fp16, const float4* in, float4* out,
                             hipStream_t stream) {
    const AtrousGrid g = atrous_grid<PG>(P.width, P.height, (uint32_t)PG);
    if (fp16)
        hipLaunchKernelGGL((k_atrous_staged<GuidesF16, PG>), dim3(g.launched), dim3(BLOCK_THREADS), 0, stream, P, g,
                           static_cast<const ushort4*>(guides), in, out);
    else
        hipLaunchKernelGGL((k_atrous_staged<GuidesF32, PG>), dim3(g.launched), dim3(BLOCK_THREADS), 0, stream, P, g,
                           static_cast<const float4*>(guides), in, out);
}

void launch_atrous(const Params& P, int step, const void* guides, bool fp16, const float4* in, float4* out,
                   hipStream_t stream) {
    if (step == 4) return launch_atrous_pg<4>(P, guides, fp16, in, out, stream);
    if (step == 2) return launch_atrous_pg<2>(P, guides, fp16, in, out, stream);
    if (step == 1) return launch_atrous_pg<1>(P, guides, fp16, in, out, stream);
    dim3 grid(stencil_blocks(P.width, P.height));
    if (fp16)
        hipLaunchKernelGGL(k_atrous<GuidesF16>, grid, dim3(BLOCK_THREADS), 0, stream, P, step,
                           static_cast<const ushort4*>(guides), in, out);
    else
        hipLaunchKernelGGL(k_atrous<GuidesF32>, grid, dim3(BLOCK_THREADS), 0, stream, P, step,
                           static_cast<const float4*>(guides), in, out);
}

void launch_display(const Params& P, const float4* accum, const float4* denoised, uchar4* out, hipStream_t stream) {
    uint32_t n = P.width * P.height;
    hipLaunchKernelGGL(k_display, dim3((n + BLOCK_THREADS - 1) / BLOCK_THREADS), dim3(BLOCK_THREADS), 0, stream, P,
                       accum, denoised, out);
}

void launch_compose(const Params& P, const float4* albedo, const float4* in, float4* out, hipStream_t stream) {
    dim3 grid((P.width + 15) / 16, (P.height + 15) / 16);
    hipLaunchKernelGGL(k_compose, grid, dim3(BLOCK_THREADS), 0, stream, P, albedo, in, out);
}

template <class G>
static void launch_atrous_batch(const DenoiseTable& D, dim3 grid, hipStream_t stream) {
    hipLaunchKernelGGL((k_atrous_batch<G, 1>), grid, dim3(BLOCK_THREADS), 0, stream, D, 1, 0);
    hipLaunchKernelGGL((k_atrous_batch<G, 2>), grid, dim3(BLOCK_THREADS), 0, stream, D, 2, 1);
    hipLaunchKernelGGL((k_atrous_batch<G, 4>), grid, dim3(BLOCK_THREADS), 0, stream, D, 4, 0);
    hipLaunchKernelGGL((k_atrous_batch<G, 0>), grid, dim3(BLOCK_THREADS), 0, stream, D, 8, 1);
    hipLaunchKernelGGL((k_atrous_batch<G, 0>), grid, dim3(BLOCK_THREADS), 0, stream, D, 16, 0);
}

void launch_denoise_batch(const DenoiseTable& D, bool fp16, hipStream_t stream) {
    const dim3 grid(D.begin[D.n]);
    if (grid.x == 0) return;
    if (fp16) launch_atrous_batch<GuidesF16>(D, grid, stream);
    else launch_atrous_batch<GuidesF32>(D, grid, stream);
    hipLaunchKernelGGL(k_compose_batch, grid, dim3(BLOCK_THREADS), 0, stream, D);
}

void launch_fill_u32(uint32_t* p, uint32_t value, size_t n, hipStream_t stream) {
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks == 0) return;
    hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)blocks), dim3(256), 0, stream, p, value, n);
}

}  // namespace rm
