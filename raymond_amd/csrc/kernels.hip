// Raymond's frame kernels for gfx950 (one wave64 per 8×8 pixel tile):
//   k_raycast        closest-hit / any-hit hit records for a batch of rays
//   k_main           samples_per_frame paths per pixel: NEE+MIS path tracing
//                    with volumes, dispersion, Russian roulette; guides
//   k_light_trace    photons from the lights through mirrors and glass,
//                    splatted to the camera and deposited in the hash grid
//   k_resolve        fixed-point splats into the accumulation
//   k_photon_gather  camera paths through delta surfaces gathering photons
//   k_prepare / k_atrous / k_compose   the à-trous denoiser
// Each kernel restates the reference entry point cited beside it
// (src/shaders.mjs). Traversal stacks live in LDS with a global spill.
#include <hip/hip_fp16.h>

#include <type_traits>

#include "kernels.hpp"

namespace rm {

// ------------------------------------------------------------------ thread mapping
// A 256-thread block covers a 16×16 pixel tile as four 8×8 wave tiles, so a
// wave's primary rays are spatially coherent.
RM_D void tile_pixel_of(uint32_t t, uint32_t& x, uint32_t& y) {
    uint32_t w = t >> 6, l = t & 63u;
    x = blockIdx.x * 16u + (w & 1u) * 8u + (l & 7u);
    y = blockIdx.y * 16u + (w >> 1) * 8u + (l >> 3);
}
RM_D void tile_pixel(uint32_t& x, uint32_t& y) { tile_pixel_of(threadIdx.x, x, y); }

// Stencil passes (the denoiser) read a neighbourhood up to 32 pixels away.
// Workgroups are dealt to the 8 XCDs round-robin by linear block id, and each
// XCD has its own L2, so a plain row-major grid makes every L2 fetch nearly
// the whole image. Instead the 16×16 tiles are ordered in vertical strips
// STRIP_TILES wide (row-major inside a strip), and each XCD takes one
// contiguous run of that order: its in-flight blocks then share a compact
// footprint that fits its L2. Launch with stencil_blocks() blocks.
constexpr uint32_t XCDS = 8u;
constexpr uint32_t STRIP_TILES = 16u;
__host__ __device__ inline uint32_t stencil_blocks(uint32_t width, uint32_t height) {
    uint32_t tiles = ((width + 15u) / 16u) * ((height + 15u) / 16u);
    return (tiles + XCDS - 1u) / XCDS * XCDS;
}
RM_D bool stencil_pixel(uint32_t width, uint32_t height, uint32_t& x, uint32_t& y) {
    const uint32_t tiles_x = (width + 15u) / 16u, tiles_y = (height + 15u) / 16u;
    const uint32_t tiles = tiles_x * tiles_y;
    const uint32_t per_xcd = (tiles + XCDS - 1u) / XCDS;
    const uint32_t t = (blockIdx.x % XCDS) * per_xcd + blockIdx.x / XCDS;
    if (t >= tiles) return false;
    const uint32_t strip = t / (STRIP_TILES * tiles_y);
    const uint32_t r = t - strip * STRIP_TILES * tiles_y;
    const uint32_t w = min(STRIP_TILES, tiles_x - strip * STRIP_TILES);
    const uint32_t ty = r / w, tx = strip * STRIP_TILES + (r - ty * w);
    const uint32_t th = threadIdx.x, wave = th >> 6, l = th & 63u;
    x = tx * 16u + (wave & 1u) * 8u + (l & 7u);
    y = ty * 16u + (wave >> 1) * 8u + (l >> 3);
    return x < width && y < height;
}

template <int LDS_DEPTH, int TREE = TREE_FP32>
struct BlockStack {
    using S = Stack<LDS_DEPTH, BLOCK_THREADS, TREE>;
    RM_D static S make(uint2* lds, uint32_t* spill_base, uint32_t slot, uint32_t slots) {
        S s;
        s.lds = lds + threadIdx.x;
        s.spill = reinterpret_cast<uint2*>(spill_base) + slot;
        s.stride = slots;
        return s;
    }
};

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

// Items (pixels) a wave of the persistent path pass claims per atomic.
constexpr uint32_t CLAIM_ITEMS = 256u;

// ------------------------------------------------------------------ radiance
struct Guides { V3 albedo; float4 normal; };

// A camera path between bounces: radiance()'s loop variables
// (src/shaders.mjs:1179-1205), so the loop can also run one bounce at a time.
struct PathState {
    V3 o, d, throughput, result, last_p;
    float last_pdf, travelled, wavelength, spectrum_u;
    uint32_t depth;
    bool last_delta, guided, caustic, diffuse_seen;
};

// Seeds a sample's RNG streams and starts its camera path: main()'s per-sample
// setup and radiance()'s prologue (src/shaders.mjs:1852-1865, 1181-1204).
RM_D void path_begin(const Params& P, Rng& g, PathState& ps, Guides& gd, uint32_t x, uint32_t y, uint32_t index,
                     uint32_t sample) {
    g.state = pcg(index ^ pcg(sample));
    g.qmc = true;
    g.seed = pcg(index + 0x9e3779b9u);
    g.index = sample;
    g.dim = 0u;
    ps.spectrum_u = (float)((pcg(index * 7919u + 1u) + sample * 2654435769u) >> 8u) * (1.0f / 16777216.0f);
    V2 j = g.rand2();
    gd.albedo = {1.0f, 1.0f, 1.0f};
    gd.normal = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    ps.guided = false;
    ps.travelled = 0.0f;
    Ray ray = camera_ray(P, g, {(float)x + j.x, (float)y + j.y}, true);
    ps.o = ray.o;
    ps.d = ray.d;
    ps.throughput = {1.0f, 1.0f, 1.0f};
    ps.result = {0.0f, 0.0f, 0.0f};
    ps.last_pdf = 0.0f;
    ps.last_delta = true;
    ps.last_p = ray.o;
    ps.wavelength = 0.0f;
    ps.caustic = false;
    ps.diffuse_seen = false;
    ps.depth = 0u;
}

// A next-event shadow ray waiting to be traced, and what it adds to the
// path's radiance when nothing blocks it: vec · (phase · T · mis / pdf) in a
// medium, vec · (T · mis / pdf) at a surface, T being the volumes'
// transmittance along the ray (1 without volumes). The factors are radiance()'s
// own, in its order (src/shaders.mjs:1247-1262, 1311-1326), split at the trace.
struct PendingNee {
    V3 o, d, vec;
    float reach, phase, mis, pdf;
    bool active, medium;
};

template <bool VOLUMES>
RM_D void apply_nee(const Ctx& C, PathState& ps, const PendingNee& pn) {
    float tr = VOLUMES ? transmittance(C, pn.o, pn.d, pn.reach) : 1.0f;
    float k = pn.medium ? pn.phase * tr * pn.mis / pn.pdf : tr * pn.mis / pn.pdf;
    ps.result = ps.result + pn.vec * k;
}

// One pass of radiance()'s bounce loop (src/shaders.mjs:1206-1356) at
// ps.depth, after its closest hit `hit` has been traced: the next-event
// shadow ray, if any, is left in pn for the caller to trace (and apply_nee)
// before the next bounce. Returns true when the path has ended (the loop's
// break, or its last bounce); ps.result then holds the sample once pn is
// applied.
template <bool VOLUMES, bool TEXTURES, bool ENV, bool DISPERSION>
RM_D bool path_shade(const Ctx& C, Rng& g, PathState& ps, Guides& gd, const Hit& hit, PendingNee& pn) {
    const Params& P = *C.P;
    const DevScene& S = C.S;
    const uint32_t choices = light_choices<ENV>(C);
    const uint32_t depth = ps.depth;
    g.dim = (depth + 1u) << 16u;
    pn.active = false;

    if (VOLUMES && P.volume_count > 0u) {
        Scatter sc = scatter_in_volumes(C, g, ps.o, ps.d, hit.t);
        if (sc.volume != NONE) {
            Mat vm = load_mat(S, __float_as_uint(S.prims[(size_t)sc.volume * 4].y));
            V3 p = ps.o + ps.d * sc.t;
            ps.travelled += sc.t;
            if (!ps.guided) {
                ps.guided = true;
                gd.albedo = ps.throughput * vm.color;
                gd.normal = make_float4(-ps.d.x, -ps.d.y, -ps.d.z, ps.travelled);
            }
            if (depth == P.max_bounces) return true;
            ps.throughput = ps.throughput * vm.color;
            if (choices > 0u) {
                LightSample ls = sample_light<ENV, TEXTURES>(C, g, p);
                if (ls.pdf > 0.0f) {
                    float phase = henyey_greenstein(dot(ps.d, ls.wi), vm.anisotropy);
                    pn.o = p;
                    pn.d = ls.wi;
                    pn.reach = ls.dist * (float)(1.0 - 1e-3);
                    pn.vec = ps.throughput * ls.radiance;
                    pn.phase = phase;
                    pn.mis = power_heuristic(ls.pdf, phase);
                    pn.pdf = ls.pdf;
                    pn.medium = true;
                    pn.active = true;
                }
            }
            V3 wi = sample_henyey_greenstein(g, ps.d, vm.anisotropy);
            ps.last_pdf = henyey_greenstein(dot(ps.d, wi), vm.anisotropy);
            ps.last_delta = false;
            ps.last_p = p;
            ps.o = p;
            ps.d = wi;
            ps.caustic = false;
            ps.diffuse_seen = true;
            if (!roulette(g, ps.throughput, depth)) return true;
            ps.depth = depth + 1u;
            return ps.depth > P.max_bounces;
        }
    }

    if (hit.prim == NONE) {
        if (!ps.guided) gd.albedo = ps.throughput;
        float sky_weight = 1.0f;
        if (has_env<ENV>(C) && !ps.last_delta)
            sky_weight = power_heuristic(ps.last_pdf, env_pdf(C, ps.d) / (float)choices);
        ps.result = ps.result + ps.throughput * sky<ENV>(C, ps.d) * sky_weight;
        if (in_sun(P, ps.d) && !(ps.caustic && ps.last_delta)) {
            float w = 1.0f;
            if (!ps.last_delta) w = power_heuristic(ps.last_pdf, 1.0f / ((float)choices * 2.0f * PI * P.sun_omc));
            ps.result = ps.result + ps.throughput * pv3(P.sun_radiance) * w;
        }
        return true;
    }

    Surface s = surface<TEXTURES>(S, ps.o, ps.d, hit);
    Mat m = load_mat(S, s.material);
    ps.travelled += hit.t;

    if (!ps.guided && (m.kind == LIGHT || !is_delta(m) || depth == P.max_bounces)) {
        ps.guided = true;
        V3 a = m.kind == LIGHT ? v3(1.0f, 1.0f, 1.0f) : albedo<TEXTURES>(S, m, s.p, s.uv);
        gd.albedo = ps.throughput * a;
        gd.normal = make_float4(s.n.x, s.n.y, s.n.z, ps.travelled);
    }

    bool glows = m.emission.x > 0.0f || m.emission.y > 0.0f || m.emission.z > 0.0f;
    if (s.front && glows && !(ps.caustic && ps.last_delta && s.light != NONE)) {
        float w = 1.0f;
        if (!ps.last_delta && s.light != NONE)
            w = power_heuristic(ps.last_pdf, light_pdf<ENV>(C, hit.prim, ps.last_p, ps.d, hit.t));
        ps.result = ps.result + ps.throughput * emitted<TEXTURES>(S, m, s.uv) * w;
    }

    if (m.kind == LIGHT) return true;
    bool delta = is_delta(m);
    if (!delta) {
        ps.caustic = P.caustics != 0u && !ps.diffuse_seen;
        ps.diffuse_seen = true;
    }
    if (depth == P.max_bounces) return true;

    V3 base = albedo<TEXTURES>(S, m, s.p, s.uv);
    V3 wo = -ps.d;

    if (!delta && choices > 0u) {
        LightSample ls = sample_light<ENV, TEXTURES>(C, g, s.p);
        if (ls.pdf > 0.0f && dot(ls.wi, s.ng) > 0.0f) {
            Eval e = eval_bsdf(m, base, s.n, wo, ls.wi);
            V3 origin = offset(s, ls.wi);
            if (e.pdf > 0.0f) {
                pn.o = origin;
                pn.d = ls.wi;
                pn.reach = (ls.dist - dot(origin - s.p, ls.wi)) * (float)(1.0 - 1e-3);
                pn.vec = ps.throughput * e.f * ls.radiance;
                pn.phase = 1.0f;
                pn.mis = power_heuristic(ls.pdf, e.pdf);
                pn.pdf = ls.pdf;
                pn.medium = false;
                pn.active = true;
            }
        }
    }

    m.ior = dispersed_ior<DISPERSION>(m, ps.wavelength, ps.throughput, ps.spectrum_u);
    BsdfSample bs = sample_bsdf(g, m, base, s, wo);
    if (!bs.ok) return true;
    ps.throughput = ps.throughput * bs.weight;
    ps.last_pdf = bs.pdf;
    ps.last_delta = bs.delta;
    ps.last_p = s.p;
    ps.o = offset(s, bs.wi);
    ps.d = bs.wi;
    if (!roulette(g, ps.throughput, depth)) return true;
    ps.depth = depth + 1u;
    return ps.depth > P.max_bounces;
}

// One bounce with its two traces in sequence: the closest hit, the shading,
// then the shadow ray. Returns true when the path has ended.
template <bool VOLUMES, bool TEXTURES, bool ENV, bool DISPERSION, class STACK>
RM_D bool path_step(const Ctx& C, Rng& g, STACK& stack, PathState& ps, Guides& gd) {
    NoCount nc;
    Hit hit = trace<false>(C.S, ps.o, ps.d, INF, stack, nc);
    PendingNee pn;
    bool ended = path_shade<VOLUMES, TEXTURES, ENV, DISPERSION>(C, g, ps, gd, hit, pn);
    if (pn.active && trace<true>(C.S, pn.o, pn.d, pn.reach, stack, nc).prim == NONE) apply_nee<VOLUMES>(C, ps, pn);
    return ended;
}

// The albedo, normal or depth seen through a pixel centre (src/shaders.mjs:1377-1407).
template <bool TEXTURES, bool ENV, class STACK>
RM_D V3 inspect(const Ctx& C, Rng& g, STACK& stack, V2 px) {
    const Params& P = *C.P;
    NoCount nc;
    Ray ray = camera_ray(P, g, px, false);
    Hit hit = trace<false>(C.S, ray.o, ray.d, INF, stack, nc);
    if (hit.prim == NONE) return P.view == VIEW_ALBEDO ? sky<ENV>(C, ray.d) : v3(0.0f, 0.0f, 0.0f);
    Surface s = surface<TEXTURES>(C.S, ray.o, ray.d, hit);
    Mat m = load_mat(C.S, s.material);
    if (P.view == VIEW_ALBEDO) {
        if (m.kind == LIGHT) {
            V3 e = emitted<TEXTURES>(C.S, m, s.uv);
            return e / fmaxf(fmaxf(e.x, e.y), fmaxf(e.z, 1.0f));
        }
        return albedo<TEXTURES>(C.S, m, s.p, s.uv);
    }
    if (P.view == VIEW_NORMAL) return s.n;
    return s3(hit.t);
}

// A pixel's sums over the frame's samples (src/shaders.mjs:1846-1894).
// sum += v with the rounding error added to err: Knuth's TwoSum, exact in
// IEEE FP32 arithmetic without contraction (this file is compiled with
// -ffp-contract=off), so sum + err carries the true running sum to about one
// rounding of the total however many frames are added.
RM_D float two_sum(float& sum, float v) {
    const float t = sum + v;
    const float z = t - sum;
    const float e = (sum - (t - z)) + (v - z);
    sum = t;
    return e;
}
RM_D void compensated_add(float4& sum_ref, float4& err_ref, float4 v) {
    float4 s = sum_ref, e = err_ref;
    e.x += two_sum(s.x, v.x);
    e.y += two_sum(s.y, v.y);
    e.z += two_sum(s.z, v.z);
    e.w += two_sum(s.w, v.w);
    sum_ref = s;
    err_ref = e;
}

struct PixelSums {
    V3 c, albedo;
    float w;
    float4 normal;
    RM_D void clear() {
        c = {0, 0, 0};
        albedo = {0, 0, 0};
        w = 0.0f;
        normal = make_float4(0, 0, 0, 0);
    }
    // Adds one sample, dropped whole if anything in it overflowed.
    RM_D void add(const Params& P, V3 v, const Guides& gd, bool clamp) {
        if (clamp) {
            float peak = fmaxf(v.x, fmaxf(v.y, v.z));
            if (P.firefly_clamp > 0.0f && peak > P.firefly_clamp) v = v * (P.firefly_clamp / peak);
        }
        if (finite3(v) && finite3(gd.albedo) && finite3(v3(gd.normal.x, gd.normal.y, gd.normal.z)) &&
            finite3(s3(gd.normal.w))) {
            float l = luminance(v);
            c = c + v;
            w += l * l;
            albedo = albedo + gd.albedo;
            normal = make_float4(normal.x + gd.normal.x, normal.y + gd.normal.y, normal.z + gd.normal.z,
                                 normal.w + gd.normal.w);
        }
    }
    // The first frame (sample offset 0) overwrites; later ones add. With
    // comp (3 float4 a pixel), each add is compensated: comp gathers the
    // rounding error of every add to the running sums.
    RM_D void store(const Params& P, uint32_t index, float4* accum, float4* galb, float4* gnorm,
                    float4* comp) const {
        const float4 add_a = make_float4(c.x, c.y, c.z, w);
        const float4 add_g = make_float4(albedo.x, albedo.y, albedo.z, 0.0f);
        if (P.sample_offset == 0u) {
            accum[index] = add_a;
            galb[index] = add_g;
            gnorm[index] = normal;
            if (comp) {
                const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                comp[(size_t)index * 3] = zero;
                comp[(size_t)index * 3 + 1] = zero;
                comp[(size_t)index * 3 + 2] = zero;
            }
        } else if (!comp) {
            float4 a = accum[index], ga = galb[index], gn = gnorm[index];
            accum[index] = make_float4(a.x + c.x, a.y + c.y, a.z + c.z, a.w + w);
            galb[index] = make_float4(ga.x + albedo.x, ga.y + albedo.y, ga.z + albedo.z, ga.w + 0.0f);
            gnorm[index] = make_float4(gn.x + normal.x, gn.y + normal.y, gn.z + normal.z, gn.w + normal.w);
        } else {
            compensated_add(accum[index], comp[(size_t)index * 3], add_a);
            compensated_add(galb[index], comp[(size_t)index * 3 + 1], add_g);
            compensated_add(gnorm[index], comp[(size_t)index * 3 + 2], normal);
        }
    }
};

// ------------------------------------------------------------------ main
// src/shaders.mjs:1838-1895, one thread per pixel: the frame's samples one
// after another, each path run to its end.
//
// Four waves per SIMD: left to itself the compiler keeps every value of a
// path in registers (153 VGPRs, three waves); capping it at 128 spills about
// 112 bytes a lane of rarely touched state (the pixel's sums, the guides) and
// the extra wave hides enough traversal latency to make the pass about 11%
// faster on the 1080p sphere scene.
// One pixel's samples of a frame: k_main's body, shared with k_main_batch.
template <bool VOLUMES, bool TEXTURES, bool ENV, bool DISPERSION, int TREE>
RM_D void main_pixel(const DevScene& S, const Params& P, uint32_t x, uint32_t y, float4* accum, float4* guide_albedo,
                     float4* guide_normal, float4* comp, uint32_t* spill, uint32_t slots, uint2* lds) {
    if (x >= P.width || y >= P.height) return;
    const Ctx C{S, &P};
    const uint32_t index = y * P.width + x;
    auto stack = BlockStack<MAIN_LDS_DEPTH, TREE>::make(lds, spill, index, slots);
    PixelSums sums;
    sums.clear();
    for (uint32_t k = 0; k < P.samples_per_frame; k++) {
        Rng g;
        PathState ps;
        Guides gd;
        if (P.view == VIEW_COLOR) {
            path_begin(P, g, ps, gd, x, y, index, P.sample_offset + k);
            while (!path_step<VOLUMES, TEXTURES, ENV, DISPERSION>(C, g, stack, ps, gd)) {
            }
            sums.add(P, ps.result, gd, true);
        } else {
            gd.albedo = {0, 0, 0};
            gd.normal = make_float4(0, 0, 0, 0);
            sums.add(P, inspect<TEXTURES, ENV>(C, g, stack, {(float)x + 0.5f, (float)y + 0.5f}), gd, false);
        }
    }
    sums.store(P, index, accum, guide_albedo, guide_normal, comp);
}

template <bool VOLUMES, bool TEXTURES, bool ENV, bool DISPERSION, int TREE>
__global__ __launch_bounds__(BLOCK_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_main(
    DevScene S, Params P, float4* __restrict__ accum, float4* __restrict__ guide_albedo,
    float4* __restrict__ guide_normal, float4* __restrict__ comp, uint32_t* spill, uint32_t slots) {
    __shared__ uint2 lds[MAIN_LDS_DEPTH * BLOCK_THREADS];
    uint32_t x, y;
    tile_pixel(x, y);
    main_pixel<VOLUMES, TEXTURES, ENV, DISPERSION, TREE>(S, P, x, y, accum, guide_albedo, guide_normal, comp, spill,
                                                         slots, lds);
}

// Several jobs' path passes in one launch (rm_frame_batch): the grid's 16×16
// tiles are the jobs' tiles one job after another, and each block renders
// its tile of its job exactly as k_main would. The job table is the kernel
// argument, so a block's job (the same for all its lanes) is read by scalar
// loads from the argument segment.
template <bool VOLUMES, bool TEXTURES, bool ENV, bool DISPERSION, int TREE>
__global__ __launch_bounds__(BLOCK_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_main_batch(BatchTable T) {
    __shared__ uint2 lds[MAIN_LDS_DEPTH * BLOCK_THREADS];
    uint32_t j = 0;
    while (j + 1u < T.n && blockIdx.x >= T.begin[j + 1u]) ++j;
    const BatchJob& J = T.job[j];
    const uint32_t tile = blockIdx.x - T.begin[j];
    const uint32_t w = threadIdx.x >> 6, l = threadIdx.x & 63u;
    const uint32_t x = (tile % J.tiles_x) * 16u + (w & 1u) * 8u + (l & 7u);
    const uint32_t y = (tile / J.tiles_x) * 16u + (w >> 1) * 8u + (l >> 3);
    main_pixel<VOLUMES, TEXTURES, ENV, DISPERSION, TREE>(J.S, J.P, x, y, J.accum, J.galb, J.gnorm, J.comp, J.spill,
                                                         J.slots, lds);
}

// The same samples with the block's live paths packed together between
// bounces. A block's 256 paths start together (coherent primary rays); after
// each bounce, when fewer waves could hold the paths still alive, their state
// moves through LDS to the first lanes of the block, so emptied waves wait at
// the barrier instead of running the bounce with a few live lanes. A path
// that ends leaves its sample in a per-lane staging record; once every path of
// the block has ended, each pixel's own thread adds its sample in order. Every
// path runs exactly the same arithmetic as in k_main, so the sums are
// bit-identical; only the colour view runs here. The barrier after every
// bounce ties a block's waves to its slowest one, so with short paths (1 spp
// per frame on the sphere scene) k_main is faster.
constexpr int COMPACT_LDS_DEPTH = 4;
constexpr int SLOT_WORDS = 27;

struct SlotStore {
    uint32_t* w;      // [SLOT_WORDS][stride]
    uint32_t stride;  // slots: BLOCK_THREADS in LDS, the path count in global memory
    RM_D void put(uint32_t slot, int k, float v) { w[(size_t)k * stride + slot] = __float_as_uint(v); }
    RM_D void putu(uint32_t slot, int k, uint32_t v) { w[(size_t)k * stride + slot] = v; }
    RM_D float get(uint32_t slot, int k) const { return __uint_as_float(w[(size_t)k * stride + slot]); }
    RM_D uint32_t getu(uint32_t slot, int k) const { return w[(size_t)k * stride + slot]; }
    RM_D void put3(uint32_t slot, int k, V3 v) { put(slot, k, v.x); put(slot, k + 1, v.y); put(slot, k + 2, v.z); }
    RM_D V3 get3(uint32_t slot, int k) const { return {get(slot, k), get(slot, k + 1), get(slot, k + 2)}; }
};

// A live path's state between bounces. The RNG's seed, sample index and
// spectrum offset are functions of the pixel and sample, and its dimension is
// reset at every bounce, so they are recomputed rather than moved.
RM_D void slot_save(SlotStore& st, uint32_t slot, const PathState& ps, const Guides& gd, uint32_t pix) {
    st.put3(slot, 0, ps.o);
    st.put3(slot, 3, ps.d);
    st.put3(slot, 6, ps.throughput);
    st.put3(slot, 9, ps.result);
    st.put3(slot, 12, ps.last_p);
    st.put(slot, 15, ps.last_pdf);
    st.put(slot, 16, ps.travelled);
    st.put(slot, 17, ps.wavelength);
    st.putu(slot, 18, ps.depth | (ps.last_delta ? 1u << 24 : 0u) | (ps.guided ? 1u << 25 : 0u) |
                          (ps.caustic ? 1u << 26 : 0u) | (ps.diffuse_seen ? 1u << 27 : 0u));
    st.put3(slot, 19, gd.albedo);
    st.put(slot, 22, gd.normal.x);
    st.put(slot, 23, gd.normal.y);
    st.put(slot, 24, gd.normal.z);
    st.put(slot, 25, gd.normal.w);
    st.putu(slot, 26, pix);
}
// The RNG and spectrum offset of pixel `index`'s sample, as path_begin seeds
// them, with the dimension back at 0 for the next bounce.
RM_D void reseed(Rng& g, PathState& ps, uint32_t index, uint32_t sample) {
    g.state = pcg(index ^ pcg(sample));
    g.qmc = true;
    g.seed = pcg(index + 0x9e3779b9u);
    g.index = sample;
    g.dim = 0u;
    ps.spectrum_u = (float)((pcg(index * 7919u + 1u) + sample * 2654435769u) >> 8u) * (1.0f / 16777216.0f);
}
// The fields slot_save wrote.
RM_D void slot_fields(const SlotStore& st, uint32_t slot, PathState& ps, Guides& gd, uint32_t& pix) {
    ps.o = st.get3(slot, 0);
    ps.d = st.get3(slot, 3);
    ps.throughput = st.get3(slot, 6);
    ps.result = st.get3(slot, 9);
    ps.last_p = st.get3(slot, 12);
    ps.last_pdf = st.get(slot, 15);
    ps.travelled = st.get(slot, 16);
    ps.wavelength = st.get(slot, 17);
    uint32_t bits = st.getu(slot, 18);
    ps.depth = bits & 0xffffffu;
    ps.last_delta = (bits >> 24) & 1u;
    ps.guided = (bits >> 25) & 1u;
    ps.caustic = (bits >> 26) & 1u;
    ps.diffuse_seen = (bits >> 27) & 1u;
    gd.albedo = st.get3(slot, 19);
    gd.normal = make_float4(st.get(slot, 22), st.get(slot, 23), st.get(slot, 24), st.get(slot, 25));
    pix = st.getu(slot, 26);
}
RM_D void slot_load(const SlotStore& st, uint32_t slot, const Params& P, uint32_t sample, PathState& ps, Guides& gd,
                    Rng& g, uint32_t& pix) {
    slot_fields(st, slot, ps, gd, pix);
    uint32_t x, y;
    tile_pixel_of(pix, x, y);
    reseed(g, ps, y * P.width + x, sample);
}
// A wavefront slot: its word 26 is the pixel's index itself.
RM_D void slot_load_wf(const SlotStore& st, uint32_t slot, uint32_t sample, PathState& ps, Guides& gd, Rng& g,
                       uint32_t& pix) {
    slot_fields(st, slot, ps, gd, pix);
    reseed(g, ps, pix, sample);
}

template <bool VOLUMES, bool TEXTURES, bool ENV, bool DISPERSION>
__global__ __launch_bounds__(BLOCK_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_main_compact(
    DevScene S, Params P, float4* __restrict__ accum, float4* __restrict__ guide_albedo,
    float4* __restrict__ guide_normal, float4* __restrict__ comp, uint32_t* spill, uint32_t slots,
    float4* __restrict__ staging) {
    __shared__ uint2 lds[COMPACT_LDS_DEPTH * BLOCK_THREADS];
    __shared__ uint32_t slot_words[SLOT_WORDS * BLOCK_THREADS];
    __shared__ uint32_t wave_live[BLOCK_THREADS / 64];
    const Ctx C{S, &P};
    const uint32_t t = threadIdx.x, wave = t >> 6, lane = t & 63u;
    const uint32_t gid = (blockIdx.y * gridDim.x + blockIdx.x) * BLOCK_THREADS + t;
    uint32_t x, y;
    tile_pixel(x, y);
    const bool valid = x < P.width && y < P.height;
    const uint32_t index = valid ? y * P.width + x : 0u;
    auto stack = BlockStack<COMPACT_LDS_DEPTH>::make(lds, spill, gid, slots);
    SlotStore st{slot_words, BLOCK_THREADS};
    float4* const stage = staging + (size_t)(gid - t) * 3;  // this block's 256 records of 3 float4
    PixelSums sums;
    sums.clear();
    for (uint32_t k = 0; k < P.samples_per_frame; k++) {
        const uint32_t sample = P.sample_offset + k;
        Rng g = {};
        PathState ps = {};
        Guides gd = {};
        uint32_t pix = t;
        bool live = valid;
        if (live) path_begin(P, g, ps, gd, x, y, index, sample);
            for (;;) {
            if (live && path_step<VOLUMES, TEXTURES, ENV, DISPERSION>(C, g, stack, ps, gd)) {
                stage[pix * 3] = make_float4(ps.result.x, ps.result.y, ps.result.z, 0.0f);
                stage[pix * 3 + 1] = make_float4(gd.albedo.x, gd.albedo.y, gd.albedo.z, 0.0f);
                stage[pix * 3 + 2] = gd.normal;
                live = false;
            }
            const uint64_t mask = __ballot(live);
            if (lane == 0) wave_live[wave] = (uint32_t)__popcll(mask);
            __syncthreads();
            uint32_t total = 0, before = 0, nonempty = 0;
            for (uint32_t w = 0; w < BLOCK_THREADS / 64; w++) {
                uint32_t n = wave_live[w];
                before += w < wave ? n : 0u;
                total += n;
                nonempty += n > 0u ? 1u : 0u;
            }
            if (total == 0u) break;
            if ((total + 63u) / 64u < nonempty) {
                // Pack the live paths into lanes 0 .. total-1, in lane order.
                if (live) slot_save(st, before + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull)), ps, gd, pix);
                __syncthreads();
                live = t < total;
                if (live) slot_load(st, t, P, sample, ps, gd, g, pix);
            }
            __syncthreads();
        }
        __syncthreads();
        if (valid) {
            float4 r = stage[t * 3], a = stage[t * 3 + 1];
            Guides own;
            own.albedo = {a.x, a.y, a.z};
            own.normal = stage[t * 3 + 2];
            sums.add(P, v3(r.x, r.y, r.z), own, true);
        }
        __syncthreads();
    }
    if (valid) sums.store(P, index, accum, guide_albedo, guide_normal, comp);
}

// The same samples with lanes kept busy: waves claim 8×8 pixel tiles from a
// global counter (one atomic per CLAIM_ITEMS pixels), and a lane whose path
// ends starts its pixel's next sample, or the next unclaimed pixel, while the
// rest of the wave keeps bouncing. A pixel's samples stay in one lane, in
// order, so every sum is bit-identical to k_main's; only the colour view runs
// here. Regeneration mixes fresh primary rays with deep secondary ones in a
// wave, which costs the primary rays' coherence: at one sample per pixel per
// frame, k_main is the faster schedule.
template <bool VOLUMES, bool TEXTURES, bool ENV, bool DISPERSION>
__global__ __launch_bounds__(BLOCK_THREADS) void k_main_persistent(DevScene S, Params P, float4* __restrict__ accum,
                                                                   float4* __restrict__ guide_albedo,
                                                                   float4* __restrict__ guide_normal,
                                                                   float4* __restrict__ comp, uint32_t* spill,
                                                                   uint32_t slots, uint32_t* counter) {
    __shared__ uint2 lds[MAIN_LDS_DEPTH * BLOCK_THREADS];
    const Ctx C{S, &P};
    const uint32_t gid = blockIdx.x * BLOCK_THREADS + threadIdx.x;
    auto stack = BlockStack<MAIN_LDS_DEPTH>::make(lds, spill, gid, slots);
    const uint32_t lane = __lane_id();
    const uint32_t tiles_x = (P.width + 7u) / 8u;
    const uint32_t items = tiles_x * ((P.height + 7u) / 8u) * 64u;

    bool alive = false;
    uint32_t x = 0, y = 0, index = 0, k = 0;
    Rng g = {};
    PathState ps = {};
    Guides gd = {};
    PixelSums sums;
    sums.clear();
    uint32_t next = 0, end = 0;  // the wave's claimed, unassigned items
    bool exhausted = false;

    for (;;) {
        uint64_t need = __ballot(!alive);
        while (need != 0 && !exhausted) {
            if (next >= end) {
                uint32_t base = 0;
                if (lane == 0) base = atomicAdd(counter, CLAIM_ITEMS);
                base = __shfl(base, 0);
                if (base >= items) {
                    exhausted = true;
                    break;
                }
                next = base;
                end = min(base + CLAIM_ITEMS, items);
            }
            uint32_t avail = end - next;
            uint32_t rank = (uint32_t)__popcll(need & ((1ull << lane) - 1ull));
            if (((need >> lane) & 1ull) && rank < avail) {
                uint32_t item = next + rank, tile = item >> 6, l = item & 63u;
                x = (tile % tiles_x) * 8u + (l & 7u);
                y = (tile / tiles_x) * 8u + (l >> 3);
                if (x < P.width && y < P.height) {
                    alive = true;
                    index = y * P.width + x;
                    k = 0;
                    sums.clear();
                    path_begin(P, g, ps, gd, x, y, index, P.sample_offset);
                }
            }
            next += min((uint32_t)__popcll(need), avail);
            need = __ballot(!alive);
        }
        if (__ballot(alive) == 0) break;
        if (alive && path_step<VOLUMES, TEXTURES, ENV, DISPERSION>(C, g, stack, ps, gd)) {
            sums.add(P, ps.result, gd, true);
            if (++k < P.samples_per_frame) {
                path_begin(P, g, ps, gd, x, y, index, P.sample_offset + k);
            } else {
                sums.store(P, index, accum, guide_albedo, guide_normal, comp);
                alive = false;
            }
        }
    }
}

// ------------------------------------------------------------------ wavefront
// RM_SCHEDULE_WAVEFRONT: a frame's camera paths advanced one bounce at a time
// for every pixel, each stage a kernel of its own, a thread per queue entry:
//
//   begin   path_begin for every pixel's sample: the first live queue;
//   extend  the closest hit of each live path (a lean traversal kernel);
//   shade   path_shade at that hit; the path goes on to the next live queue
//           or, ended, to the done queue, its shadow ray to the shadow queue;
//   shadow  the any-hit test of each shadow ray, then apply_nee to its
//           path's result wherever that path went;
//   finish  PixelSums::add for the paths that ended;
//   store   PixelSums::store once the frame's samples are all in.
//
// A path's state lives at its queue slot, double-buffered between bounces,
// so every stage reads and writes its queue in order; slots are handed out
// one atomic per wave, so a wave's writes are contiguous. The same functions
// in the same order per path as k_main, each pixel's samples added in sample
// order: the accumulation is the same, bit for bit.
constexpr int NEE_WORDS = 14;    // o, d, vec, reach, phase, mis, pdf, medium
constexpr int DONE_WORDS = 11;   // result, guide albedo, guide normal, pixel
static_assert(WF_STATE_WORDS == SLOT_WORDS && WF_NEE_WORDS == NEE_WORDS + 1 && WF_DONE_WORDS == DONE_WORDS,
              "the host sizes these buffers");

// A slot for each lane with pred, from `count`, one atomic per wave; lanes
// without pred get NONE.
RM_D uint32_t wave_slot(uint32_t* count, bool pred) {
    const uint64_t mask = __ballot(pred);
    if (mask == 0ull) return NONE;
    const uint32_t lane = __lane_id();
    const int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if ((int)lane == leader) base = atomicAdd(count, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    return pred ? base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull)) : NONE;
}

RM_D SlotStore wf_state(const Wavefront& W, uint32_t which) { return SlotStore{W.state[which], W.n}; }

RM_D PixelSums wf_sums(const Wavefront& W, uint32_t i) {
    const float4 a = W.sums[i], b = W.sums[(size_t)W.n + i], c = W.sums[(size_t)2 * W.n + i];
    PixelSums s;
    s.c = {a.x, a.y, a.z};
    s.w = a.w;
    s.albedo = {b.x, b.y, b.z};
    s.normal = c;
    return s;
}

RM_D void wf_put_sums(const Wavefront& W, uint32_t i, const PixelSums& s) {
    W.sums[i] = make_float4(s.c.x, s.c.y, s.c.z, s.w);
    W.sums[(size_t)W.n + i] = make_float4(s.albedo.x, s.albedo.y, s.albedo.z, 0.0f);
    W.sums[(size_t)2 * W.n + i] = s.normal;
}

// One thread per queue entry: every stage's grid covers all n paths, and the
// threads past the queue's length (read on the device) return at once, so
// finished waves give way to new ones as in a kernel launched per ray.
#define WF_EACH(i, count) if (const uint32_t i = blockIdx.x * BLOCK_THREADS + threadIdx.x; i < (count))

// Every pixel's sample `sample`: its camera path, in pixel order in live
// state 0; the first sample of the frame also clears the pixel's sums.
__global__ __launch_bounds__(BLOCK_THREADS) void k_wf_begin(Params P, Wavefront W, uint32_t sample, uint32_t first) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        W.counts[0] = W.n;
        W.counts[1] = W.counts[2] = W.counts[3] = 0u;
    }
    WF_EACH(i, W.n) {
        Rng g;
        PathState ps;
        Guides gd;
        path_begin(P, g, ps, gd, i % P.width, i / P.width, i, sample);
        SlotStore st = wf_state(W, 0);
        slot_save(st, i, ps, gd, i);
        if (first) {
            PixelSums s;
            s.clear();
            wf_put_sums(W, i, s);
        }
    }
}

template <int TREE>
__global__ __launch_bounds__(BLOCK_THREADS) void k_wf_extend(DevScene S, Wavefront W, uint32_t cur, uint32_t* spill,
                                                             uint32_t slots) {
    __shared__ uint2 lds[TRACE_LDS_DEPTH * BLOCK_THREADS];
    // Nothing reads these until the shade kernel after this one.
    if (blockIdx.x == 0 && threadIdx.x == 0) W.counts[cur ^ 1u] = W.counts[2] = W.counts[3] = 0u;
    WF_EACH(i, W.counts[cur]) {
        auto stack = BlockStack<TRACE_LDS_DEPTH, TREE>::make(lds, spill, i, slots);
        const SlotStore st = wf_state(W, cur);
        NoCount nc;
        const Hit h = trace<false>(S, st.get3(i, 0), st.get3(i, 3), INF, stack, nc);
        W.hits[i] = make_uint4(__float_as_uint(h.t), h.prim, __float_as_uint(h.u), __float_as_uint(h.v));
    }
}

template <bool VOLUMES, bool TEXTURES, bool ENV, bool DISPERSION>
__global__ __launch_bounds__(BLOCK_THREADS) void k_wf_shade(DevScene S, Params P, Wavefront W, uint32_t cur,
                                                            uint32_t sample) {
    const Ctx C{S, &P};
    const uint32_t next = cur ^ 1u;
    const uint32_t count = W.counts[cur];
    // Whole waves reach the slot atomics together, so their ballots see every lane.
    if (blockIdx.x * BLOCK_THREADS >= count) return;
    const uint32_t i = blockIdx.x * BLOCK_THREADS + threadIdx.x;
    const bool live = i < count;
    bool ended = false;
    PendingNee pn;
    pn.active = false;
    Rng g;
    PathState ps;
    Guides gd;
    uint32_t pix = 0;
    if (live) {
        const SlotStore st = wf_state(W, cur);
        slot_load_wf(st, i, sample, ps, gd, g, pix);
        const uint4 h = W.hits[i];
        const Hit hit = {__uint_as_float(h.x), h.y, __uint_as_float(h.z), __uint_as_float(h.w)};
        ended = path_shade<VOLUMES, TEXTURES, ENV, DISPERSION>(C, g, ps, gd, hit, pn);
    }
    const uint32_t go = wave_slot(W.counts + next, live && !ended);
    const uint32_t stop = wave_slot(W.counts + 3, live && ended);
    const uint32_t shadow = wave_slot(W.counts + 2, live && pn.active);
    if (!live) return;
    if (ended) {
        SlotStore dn{W.done, W.n};
        dn.put3(stop, 0, ps.result);
        dn.put3(stop, 3, gd.albedo);
        dn.put(stop, 6, gd.normal.x);
        dn.put(stop, 7, gd.normal.y);
        dn.put(stop, 8, gd.normal.z);
        dn.put(stop, 9, gd.normal.w);
        dn.putu(stop, 10, pix);
    } else {
        SlotStore st = wf_state(W, next);
        slot_save(st, go, ps, gd, pix);
    }
    if (pn.active) {
        SlotStore nz{W.nee, W.n};
        nz.put3(shadow, 0, pn.o);
        nz.put3(shadow, 3, pn.d);
        nz.put3(shadow, 6, pn.vec);
        nz.put(shadow, 9, pn.reach);
        nz.put(shadow, 10, pn.phase);
        nz.put(shadow, 11, pn.mis);
        nz.put(shadow, 12, pn.pdf);
        nz.putu(shadow, 13, pn.medium ? 1u : 0u);
        // Where the path's result went: its done slot (top bit) or its next live slot.
        nz.putu(shadow, 14, ended ? (stop | 0x80000000u) : go);
    }
}

template <bool VOLUMES, int TREE>
__global__ __launch_bounds__(BLOCK_THREADS) void k_wf_shadow(DevScene S, Params P, Wavefront W, uint32_t next,
                                                             uint32_t* spill, uint32_t slots) {
    __shared__ uint2 lds[TRACE_LDS_DEPTH * BLOCK_THREADS];
    const Ctx C{S, &P};
    WF_EACH(i, W.counts[2]) {
        auto stack = BlockStack<TRACE_LDS_DEPTH, TREE>::make(lds, spill, i, slots);
        const SlotStore nz{W.nee, W.n};
        PendingNee pn;
        pn.o = nz.get3(i, 0);
        pn.d = nz.get3(i, 3);
        pn.vec = nz.get3(i, 6);
        pn.reach = nz.get(i, 9);
        pn.phase = nz.get(i, 10);
        pn.mis = nz.get(i, 11);
        pn.pdf = nz.get(i, 12);
        pn.medium = nz.getu(i, 13) != 0u;
        pn.active = true;
        NoCount nc;
        if (trace<true>(S, pn.o, pn.d, pn.reach, stack, nc).prim == NONE) {
            const uint32_t to = nz.getu(i, 14);
            SlotStore dst = (to & 0x80000000u) ? SlotStore{W.done, W.n} : wf_state(W, next);
            const uint32_t slot = to & 0x7fffffffu;
            const int word = (to & 0x80000000u) ? 0 : 9;   // the result's first word
            PathState ps = {};
            ps.result = dst.get3(slot, word);
            apply_nee<VOLUMES>(C, ps, pn);
            dst.put3(slot, word, ps.result);
        }
    }
}

__global__ __launch_bounds__(BLOCK_THREADS) void k_wf_finish(Params P, Wavefront W) {
    WF_EACH(i, W.counts[3]) {
        const SlotStore dn{W.done, W.n};
        Guides gd;
        gd.albedo = dn.get3(i, 3);
        gd.normal = make_float4(dn.get(i, 6), dn.get(i, 7), dn.get(i, 8), dn.get(i, 9));
        const uint32_t pix = dn.getu(i, 10);
        PixelSums s = wf_sums(W, pix);
        s.add(P, dn.get3(i, 0), gd, true);
        wf_put_sums(W, pix, s);
    }
}

__global__ __launch_bounds__(BLOCK_THREADS) void k_wf_store(Params P, Wavefront W, float4* __restrict__ accum,
                                                            float4* __restrict__ galb, float4* __restrict__ gnorm,
                                                            float4* __restrict__ comp) {
    WF_EACH(i, W.n) wf_sums(W, i).store(P, i, accum, galb, gnorm, comp);
}

#undef WF_EACH

// ------------------------------------------------------------------ light tracing
RM_D uint32_t photon_count(const Params& P) {
    return min(P.width * P.height * P.samples_per_frame, PHOTON_CAPACITY);
}
RM_D uint32_t photon_cell_hash(int32_t cx, int32_t cy, int32_t cz) {
    return (((uint32_t)cx * 73856093u) ^ ((uint32_t)cy * 19349663u) ^ ((uint32_t)cz * 83492791u)) & (PHOTON_TABLE - 1u);
}

// Adds a photon's light to the pixel that sees surface s (src/shaders.mjs:1444-1491).
template <class STACK>
RM_D void splat(const Ctx& C, Rng& g, STACK& stack, uint32_t* splats, const Surface& s, const Mat& m, V3 base, V3 d,
                V3 beta) {
    const Params& P = *C.P;
    NoCount nc;
    V3 to_camera = pv3(P.cam_pos) - s.p;
    float dist = length(to_camera);
    V3 wo = to_camera / dist;
    float cos_film = -dot(wo, pv3(P.cam_forward));
    if (cos_film <= 1e-6f || dot(wo, s.ng) <= 0.0f) return;
    float x = -dot(wo, pv3(P.cam_right)) / (cos_film * P.tan_half_fov * P.aspect);
    float y = -dot(wo, pv3(P.cam_up)) / (cos_film * P.tan_half_fov);
    float px = (x + 1.0f) * 0.5f * (float)P.width;
    float py = (1.0f - y) * 0.5f * (float)P.height;
    if (px < 0.0f || py < 0.0f || px >= (float)P.width || py >= (float)P.height) return;
    V3 wi = -d;
    Eval e = eval_bsdf(m, base, s.n, wo, wi);
    float cos_in = dot(s.n, wi);
    if (e.pdf <= 0.0f || cos_in <= 0.0f) return;
    V3 origin = offset(s, wo);
    if (trace<true>(C.S, origin, wo, dist - dot(origin - s.p, wo), stack, nc).prim != NONE) return;
    float pixel_area = (4.0f * P.tan_half_fov * P.tan_half_fov * P.aspect) / (float)(P.width * P.height);
    V3 value = beta * (e.f / cos_in) * dot(wo, s.ng) / (dist * dist * pixel_area * cos_film * cos_film * cos_film);
    if (!finite3(value)) return;
    uint32_t index = (to_u32(py) * P.width + to_u32(px)) * 4u;
    float r0 = g.rand(), r1 = g.rand(), r2 = g.rand();
    V3 fx = vmin(value * SPLAT_SCALE + v3(r0, r1, r2), s3(4e9f));
    atomicAdd(&splats[index], to_u32(fx.x));
    atomicAdd(&splats[index + 1u], to_u32(fx.y));
    atomicAdd(&splats[index + 2u], to_u32(fx.z));
}

// One photon per thread (src/shaders.mjs:1630-1814). Deposits happen after
// the bounce loop with one counter atomic per wave (ballot + prefix count):
// photon slots are claimed in a different order than one atomic per lane
// would give, which the reference leaves unspecified too.
template <bool TEXTURES, bool DISPERSION, int TREE>
__global__ __launch_bounds__(BLOCK_THREADS) void k_light_trace(DevScene S, Params P, const float4* __restrict__ E,
                                                               uint32_t* splats, uint32_t* heads,
                                                               float4* __restrict__ photons, uint32_t* spill,
                                                               uint32_t slots) {
    __shared__ uint2 lds[MAIN_LDS_DEPTH * BLOCK_THREADS];
    const Ctx C{S, &P};
    const uint32_t id = blockIdx.x * BLOCK_THREADS + threadIdx.x;
    const uint32_t photons_n = photon_count(P);
    auto stack = BlockStack<MAIN_LDS_DEPTH, TREE>::make(lds, spill, id, slots);
    bool pending = false;
    V3 dep_p = {0, 0, 0}, dep_d = {0, 0, 0}, dep_beta = {0, 0, 0};

    if (id < photons_n) {
        Rng g;
        g.qmc = false;
        g.state = pcg(id ^ pcg(P.sample_offset * 2654435769u + 977u));
        const float4 header = E[0];
        const uint32_t sources = __float_as_uint(header.x), targets = __float_as_uint(header.y);
        float xi = g.rand();
        uint32_t k = 0u;
        while (!(k + 1u >= sources || E[1u + k].x > xi)) k++;
        const float4 entry = E[1u + k];
        const float chance = entry.z;
        const uint32_t prim_index = __float_as_uint(entry.y);
        const uint32_t first_target = 1u + sources;
        bool alive = chance > 0.0f;
        V3 o = {0, 0, 0}, d = {0, 0, 0}, beta = {0, 0, 0};
        if (alive && prim_index == NONE) {
            d = -sample_cone(g, pv3(P.sun_dir), P.sun_omc);
            float pick = g.rand() * header.w / PI;
            uint32_t j = 0u;
            for (;;) {
                float r = E[first_target + j].w;
                if (j + 1u >= targets || pick < r * r) break;
                pick -= r * r;
                j++;
            }
            float4 circle = E[first_target + j];
            Frame f = frame(d);
            V2 disk = concentric_disk(g);
            disk = {disk.x * circle.w, disk.y * circle.w};
            o = xyz(circle) - d * header.z + f.c0 * disk.x + f.c1 * disk.y;
            float through = 1.0f;
            for (uint32_t i = 0u; i < targets; i++) {
                float4 c = E[first_target + i];
                V3 q = o - xyz(c);
                V3 across = q - d * dot(q, d);
                through += (i != j && dot(across, across) < c.w * c.w) ? 1.0f : 0.0f;
            }
            V3 irradiance = pv3(P.sun_radiance) * (TWO_PI * P.sun_omc);
            beta = irradiance * header.w / (through * chance * (float)photons_n);
        } else if (alive) {
            const float4* pr = S.prims + (size_t)prim_index * 4;
            float4 hd = pr[0], p0 = pr[1], p1 = pr[2], p2 = pr[3];
            Mat m = load_mat(S, __float_as_uint(hd.y));
            V3 n;
            float area;
            V2 uv;
            if (__float_as_uint(hd.x) == QUAD) {
                float a = g.rand();
                float b = g.rand();
                o = xyz(p0) + xyz(p1) * a + xyz(p2) * b;
                V3 c = cross(xyz(p1), xyz(p2));
                area = length(c);
                n = c / area;
                uv = {a, 1.0f - b};
            } else {
                float z = 1.0f - 2.0f * g.rand();
                float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
                float phi = TWO_PI * g.rand();
                n = {r * cosf(phi), z, r * sinf(phi)};
                o = xyz(p0) + n * p0.w;
                area = FOUR_PI * p0.w * p0.w;
                uv = sphere_uv(n, p1, p2);
            }
            float cones = 0.0f;
            bool inside = false;
            for (uint32_t i = 0u; i < targets; i++) {
                float4 aim = E[first_target + i];
                V3 to = xyz(aim) - o;
                float d2 = dot(to, to);
                inside = inside || d2 <= aim.w * aim.w;
                cones += sphere_omc(aim.w, fmaxf(d2, aim.w * aim.w));
            }
            if (!inside) {
                float pick = g.rand() * cones;
                uint32_t j = 0u;
                for (;;) {
                    float4 aim = E[first_target + j];
                    V3 to = xyz(aim) - o;
                    float omc = sphere_omc(aim.w, dot(to, to));
                    if (j + 1u >= targets || pick < omc) {
                        d = sample_cone(g, to / length(to), omc);
                        break;
                    }
                    pick -= omc;
                    j++;
                }
                float within = 1.0f;
                for (uint32_t i = 0u; i < targets; i++) {
                    float4 aim = E[first_target + i];
                    V3 to = xyz(aim) - o;
                    float omc = sphere_omc(aim.w, dot(to, to));
                    within += (i != j && 1.0f - dot(d, to / length(to)) <= omc) ? 1.0f : 0.0f;
                }
                float cos_out = dot(d, n);
                if (cos_out <= 0.0f) alive = false;
                else beta = emitted<TEXTURES>(S, m, uv) * (cos_out * area * 2.0f * PI * cones) /
                            (within * chance * (float)photons_n);
            } else {
                d = cosine_hemisphere(g, n);
                beta = emitted<TEXTURES>(S, m, uv) * (PI * area) / (chance * (float)photons_n);
            }
            o = o + n * (1e-4f * fmaxf(1.0f, fmaxf(fabsf(o.x), fmaxf(fabsf(o.y), fabsf(o.z)))));
        }

        float wavelength = 0.0f;
        uint32_t bounces = 0u;
        V3 carried = {1.0f, 1.0f, 1.0f};
        NoCount nc;
        for (uint32_t depth = 0u; alive && depth < P.max_bounces; depth++) {
            Hit hit = trace<false>(S, o, d, INF, stack, nc);
            if (hit.prim == NONE) break;
            Surface s = surface<TEXTURES>(S, o, d, hit);
            Mat m = load_mat(S, s.material);
            if (m.kind == LIGHT) break;
            V3 base = albedo<TEXTURES>(S, m, s.p, s.uv);
            if (!is_delta(m)) {
                if (bounces > 0u) {
                    splat(C, g, stack, splats, s, m, base, d, beta * carried);
                    pending = true;
                    dep_p = s.p;
                    dep_d = d;
                    dep_beta = beta * carried;
                }
                break;
            }
            float u = g.rand();
            m.ior = dispersed_ior<DISPERSION>(m, wavelength, carried, u);
            BsdfSample bs = sample_bsdf(g, m, base, s, -d);
            if (!bs.ok) break;
            carried = carried * bs.weight;
            bounces++;
            o = offset(s, bs.wi);
            d = bs.wi;
            if (!roulette(g, carried, depth)) break;
        }
    }

    // deposit() (src/shaders.mjs:1518-1528), one counter atomic per wave.
    uint64_t mask = __ballot(pending);
    if (mask == 0) return;
    uint32_t lane = __lane_id();
    uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&heads[PHOTON_TABLE], (uint32_t)__popcll(mask));
    base = __shfl(base, (int)leader);
    if (!pending) return;
    uint32_t i = base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    if (i >= PHOTON_CAPACITY) return;
    V3 q = dep_p / P.photon_cell;
    uint32_t h = photon_cell_hash(to_i32(floorf(q.x)), to_i32(floorf(q.y)), to_i32(floorf(q.z)));
    uint32_t next = atomicExch(&heads[h], i + 1u);
    photons[(size_t)i * 3] = make_float4(dep_p.x, dep_p.y, dep_p.z, __uint_as_float(next));
    photons[(size_t)i * 3 + 1] = make_float4(dep_beta.x, dep_beta.y, dep_beta.z, 0.0f);
    photons[(size_t)i * 3 + 2] = make_float4(-dep_d.x, -dep_d.y, -dep_d.z, 0.0f);
}

// resolve_splats() then photon_gather()'s sum into the accumulation
// (src/shaders.mjs:1819-1831, 1610-1624), in that order, for one pixel. The
// gather pass leaves its light in `gathered` (w = 1 where it adds any), so it
// can run while the path-tracing pass still writes the accumulation.
__global__ __launch_bounds__(BLOCK_THREADS) void k_resolve(Params P, uint4* splats, const float4* gathered,
                                                           float4* accum, float4* comp) {
    uint32_t i = blockIdx.x * BLOCK_THREADS + threadIdx.x;
    if (i >= P.width * P.height) return;
    uint4 s = splats[i];
    splats[i] = make_uint4(0u, 0u, 0u, s.w);
    V3 light = v3((float)s.x, (float)s.y, (float)s.z) / SPLAT_SCALE;
    float l = luminance(light);
    float spp = (float)P.samples_per_frame;
    const float4 splat = make_float4(light.x * spp, light.y * spp, light.z * spp, l * l * spp);
    float4 g = gathered[i];
    float4 gather = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (g.w != 0.0f) {
        V3 gl = v3(g.x, g.y, g.z);
        float lg = luminance(gl);
        gather = make_float4(gl.x * spp, gl.y * spp, gl.z * spp, lg * lg * spp);
    }
    float4 a = accum[i];
    if (comp) {
        float4 e = comp[(size_t)i * 3];
        compensated_add(a, e, splat);
        if (g.w != 0.0f) compensated_add(a, e, gather);
        comp[(size_t)i * 3] = e;
    } else {
        a = make_float4(a.x + splat.x, a.y + splat.y, a.z + splat.z, a.w + splat.w);
        if (g.w != 0.0f) a = make_float4(a.x + gather.x, a.y + gather.y, a.z + gather.z, a.w + gather.w);
    }
    accum[i] = a;
}

// The compensated sums' best FP32 value, sum + error, for their readers
// (the denoiser, readback, display): three planes of W·H float4.
__global__ __launch_bounds__(BLOCK_THREADS) void k_fold(Params P, const float4* __restrict__ accum,
                                                        const float4* __restrict__ galb,
                                                        const float4* __restrict__ gnorm,
                                                        const float4* __restrict__ comp, float4* __restrict__ out) {
    const uint32_t n = P.width * P.height;
    const uint32_t i = blockIdx.x * BLOCK_THREADS + threadIdx.x;
    if (i >= n) return;
    const float4* src[3] = {accum, galb, gnorm};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float4 s = src[k][i], e = comp[(size_t)i * 3 + k];
        out[(size_t)k * n + i] = make_float4(s.x + e.x, s.y + e.y, s.z + e.z, s.w + e.w);
    }
}

// The photons within the radius of a surface, times the BSDF, over the
// disc's area (src/shaders.mjs:1534-1570).
RM_D V3 gather(const Ctx& C, const uint32_t* heads, const float4* photons, const Surface& s, const Mat& m, V3 base,
               V3 wo) {
    const Params& P = *C.P;
    float r = P.photon_radius, r2 = r * r;
    V3 here = s.p / P.photon_cell;
    int32_t cx = to_i32(floorf(here.x)), cy = to_i32(floorf(here.y)), cz = to_i32(floorf(here.z));
    int32_t sx = (here.x - floorf(here.x)) >= 0.5f ? 1 : -1;
    int32_t sy = (here.y - floorf(here.y)) >= 0.5f ? 1 : -1;
    int32_t sz = (here.z - floorf(here.z)) >= 0.5f ? 1 : -1;
    V3 sum = {0, 0, 0};
    uint32_t visited = 0u;
    for (uint32_t k = 0u; k < 8u; k++) {
        uint32_t index = heads[photon_cell_hash(cx + (int32_t)(k & 1u) * sx, cy + (int32_t)((k >> 1u) & 1u) * sy,
                                                cz + (int32_t)((k >> 2u) & 1u) * sz)];
        while (index != 0u && visited < 4096u) {
            visited++;
            float4 at = photons[(size_t)(index - 1u) * 3];
            V3 off = xyz(at) - s.p;
            if (dot(off, off) < r2 && fabsf(dot(off, s.ng)) < 0.25f * r) {
                V3 wi = xyz(photons[(size_t)(index - 1u) * 3 + 2]);
                float cos_in = dot(s.n, wi);
                if (cos_in > 0.0f && dot(s.ng, wi) > 0.0f) {
                    Eval e = eval_bsdf(m, base, s.n, wo, wi);
                    sum = sum + xyz(photons[(size_t)(index - 1u) * 3 + 1]) * (e.f / cos_in);
                }
            }
            index = __float_as_uint(at.w);
        }
    }
    return sum / (PI * r2);
}

// src/shaders.mjs:1577-1626. The pixel's gathered light goes to
// gathered[index] (w = 1), or zeros when it adds none; k_resolve adds it.
template <bool TEXTURES, class STACK>
RM_D float4 gather_pixel(const Ctx& C, const uint32_t* heads, const float4* photons, uint32_t x, uint32_t y,
                         uint32_t index, STACK& stack) {
    const Params& P = *C.P;
    const DevScene& S = C.S;
    NoCount nc;
    Rng g;
    g.qmc = false;
    g.state = pcg(index ^ pcg(P.sample_offset * 2654435769u + 4099u));
    float jx = g.rand();
    float jy = g.rand();
    Ray ray = camera_ray(P, g, {(float)x + jx, (float)y + jy}, true);
    V3 o = ray.o, d = ray.d;
    V3 throughput = {1, 1, 1};
    for (uint32_t depth = 0u; depth < P.max_bounces; depth++) {
        Hit hit = trace<false>(S, o, d, INF, stack, nc);
        if (hit.prim == NONE) break;
        Surface s = surface<TEXTURES>(S, o, d, hit);
        Mat m = load_mat(S, s.material);
        if (m.kind == LIGHT) break;
        V3 base = albedo<TEXTURES>(S, m, s.p, s.uv);
        if (!is_delta(m)) {
            if (depth > 0u) {
                V3 light = throughput * gather(C, heads, photons, s, m, base, -d);
                if (finite3(light)) return make_float4(light.x, light.y, light.z, 1.0f);
            }
            break;
        }
        BsdfSample bs = sample_bsdf(g, m, base, s, -d);
        if (!bs.ok) break;
        throughput = throughput * bs.weight;
        o = offset(s, bs.wi);
        d = bs.wi;
    }
    return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

template <bool TEXTURES, int TREE>
__global__ __launch_bounds__(BLOCK_THREADS) void k_photon_gather(DevScene S, Params P, const uint32_t* heads,
                                                                 const float4* __restrict__ photons,
                                                                 float4* __restrict__ gathered, uint32_t* spill,
                                                                 uint32_t slots) {
    __shared__ uint2 lds[GATHER_LDS_DEPTH * BLOCK_THREADS];
    uint32_t x, y;
    tile_pixel(x, y);
    if (x >= P.width || y >= P.height) return;
    const Ctx C{S, &P};
    const uint32_t index = y * P.width + x;
    auto stack = BlockStack<GATHER_LDS_DEPTH, TREE>::make(lds, spill, index, slots);
    gathered[index] = gather_pixel<TEXTURES>(C, heads, photons, x, y, index, stack);
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
    float a = luminance(own_albedo);
    float variance;
    if (n >= 4.0f) {
        float mean = luminance(v3(ac.x, ac.y, ac.z)) / n;
        float spread = fmaxf(ac.w / n - mean * mean, 0.0f) * n / (n - 1.0f);
        variance = spread / n / (a * a);
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

// src/shaders.mjs:2023-2086. The 25 taps' weights are the reference's
// B3 × max(n·n', 0)^64 × exp(-|Δz| / (0.02 z step)) × exp(-|Δl| / σ), with
// the power by six squarings, the two exponentials merged into one exp2,
// and the centre's divisions hoisted to reciprocals: a few ulps from the
// reference's per-tap pow/exp/divide, several times fewer instructions.
template <class G>
__global__ __launch_bounds__(BLOCK_THREADS) void k_atrous(Params P, int step, const typename G::T* __restrict__ guides,
                                                          const float4* __restrict__ light_in,
                                                          float4* __restrict__ light_out) {
    uint32_t x, y;
    if (!stencil_pixel(P.width, P.height, x, y)) return;
    const uint32_t i = y * P.width + x;
    const float4 centre = light_in[i];
    const float4 gp = G::get(guides, i);
    const float lp = luminance(v3(centre.x, centre.y, centre.z));
    const int W = (int)P.width, H = (int)P.height;
    float variance = 0.0f, variance_weight = 0.0f;
    for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
            int xx = (int)x + dx, yy = (int)y + dy;
            if (xx >= 0 && yy >= 0 && xx < W && yy < H) {
                float w = kernel_w(abs(dx)) * kernel_w(abs(dy));
                variance += w * light_in[(uint32_t)yy * P.width + (uint32_t)xx].w;
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
        int yy = (int)y + dy * step;
        if (yy < 0 || yy >= H) continue;
#pragma unroll
        for (int dx = -2; dx <= 2; dx++) {
            int xx = (int)x + dx * step;
            if (xx < 0 || xx >= W) continue;
            uint32_t j = (uint32_t)yy * P.width + (uint32_t)xx;
            float4 q = light_in[j];
            float4 g = G::get(guides, j);
            float e = -fabsf(lp - luminance(v3(q.x, q.y, q.z))) * k_l;
            float s;
            if (sky || g.w <= 0.0f) {
                s = (sky && g.w <= 0.0f) ? 1.0f : 0.0f;
            } else if (no_normal) {
                s = 0.0f;
            } else {
                s = pow64(fmaxf(gp.x * g.x + gp.y * g.y + gp.z * g.z, 0.0f));
                e -= fabsf(gp.w - g.w) * k_z;
            }
            float w = kernel_w(abs(dx)) * kernel_w(abs(dy)) * s * exp2f(e);
            sum = sum + w * v3(q.x, q.y, q.z);
            sum_variance += w * w * q.w;
            weight += w;
        }
    }
    if (weight > 0.0f) {
        V3 r = sum / weight;
        light_out[i] = make_float4(r.x, r.y, r.z, sum_variance / (weight * weight));
    } else {
        light_out[i] = centre;
    }
}

// src/shaders.mjs:2088-2098, with the albedo k_prepare recorded.
__global__ __launch_bounds__(BLOCK_THREADS) void k_compose(Params P, const float4* __restrict__ albedo,
                                                           const float4* __restrict__ light_in, float4* out) {
    uint32_t x, y;
    tile_pixel(x, y);
    if (x >= P.width || y >= P.height) return;
    const uint32_t i = y * P.width + x;
    float4 l = light_in[i], a = albedo[i];
    V3 r = v3(l.x, l.y, l.z) * v3(a.x, a.y, a.z);
    out[i] = make_float4(r.x, r.y, r.z, 0.0f);
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

// Calls f with the tree kind as a compile-time constant.
template <class F>
static void with_tree(int tree, F&& f) {
    switch (tree) {
    case TREE_WIDE: f(std::integral_constant<int, TREE_WIDE>{}); break;
    case TREE_QUANTIZED: f(std::integral_constant<int, TREE_QUANTIZED>{}); break;
    default: f(std::integral_constant<int, TREE_FP32>{}); break;
    }
}

template <bool V, bool T, bool E, bool D>
static void launch_main_v(const LaunchMain& L, hipStream_t stream) {
    if (L.persistent_blocks > 0 && L.P.view == VIEW_COLOR) {
        hipLaunchKernelGGL((k_main_persistent<V, T, E, D>), dim3(L.persistent_blocks), dim3(BLOCK_THREADS), 0, stream,
                           L.S, L.P, L.accum, L.galb, L.gnorm, L.comp, L.spill, L.slots, L.counter);
        return;
    }
    dim3 grid((L.P.width + 15) / 16, (L.P.height + 15) / 16);
    if (L.staging && L.P.view == VIEW_COLOR) {
        hipLaunchKernelGGL((k_main_compact<V, T, E, D>), grid, dim3(BLOCK_THREADS), 0, stream, L.S, L.P, L.accum,
                           L.galb, L.gnorm, L.comp, L.spill, L.slots, L.staging);
        return;
    }
    with_tree(L.tree, [&](auto tree) {
        hipLaunchKernelGGL((k_main<V, T, E, D, decltype(tree)::value>), grid, dim3(BLOCK_THREADS), 0, stream, L.S, L.P,
                           L.accum, L.galb, L.gnorm, L.comp, L.spill, L.slots);
    });
}

template <bool V, bool T, bool E, bool D>
static int resident_blocks_v() {
    static int blocks = -1;
    if (blocks < 0) {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_main_persistent<V, T, E, D>, BLOCK_THREADS, 0) !=
            hipSuccess)
            n = 1;
        blocks = n > 0 ? n : 1;
    }
    return blocks;
}

int main_resident_blocks_per_cu(const Features& f) {
    int key = (f.volumes ? 8 : 0) | (f.textures ? 4 : 0) | (f.environment ? 2 : 0) | (f.dispersion ? 1 : 0);
    switch (key) {
    case 0: return resident_blocks_v<false, false, false, false>();
    case 1: return resident_blocks_v<false, false, false, true>();
    case 2: return resident_blocks_v<false, false, true, false>();
    case 3: return resident_blocks_v<false, false, true, true>();
    case 8: return resident_blocks_v<true, false, false, false>();
    case 10: return resident_blocks_v<true, false, true, false>();
    default: return resident_blocks_v<true, true, true, true>();
    }
}

void launch_main(const LaunchMain& L, const Features& f, hipStream_t stream) {
    int key = (f.volumes ? 8 : 0) | (f.textures ? 4 : 0) | (f.environment ? 2 : 0) | (f.dispersion ? 1 : 0);
    switch (key) {
    case 0: launch_main_v<false, false, false, false>(L, stream); break;
    case 1: launch_main_v<false, false, false, true>(L, stream); break;
    case 2: launch_main_v<false, false, true, false>(L, stream); break;
    case 3: launch_main_v<false, false, true, true>(L, stream); break;
    case 8: launch_main_v<true, false, false, false>(L, stream); break;
    case 10: launch_main_v<true, false, true, false>(L, stream); break;
    default: launch_main_v<true, true, true, true>(L, stream); break;
    }
}

template <bool V, bool T, bool E, bool D>
static void launch_wavefront_v(const LaunchWavefront& L, hipStream_t stream) {
    const dim3 grid(L.blocks), block(BLOCK_THREADS);
    for (uint32_t k = 0; k < L.P.samples_per_frame; ++k) {
        const uint32_t sample = L.P.sample_offset + k;
        hipLaunchKernelGGL(k_wf_begin, grid, block, 0, stream, L.P, L.W, sample, k == 0 ? 1u : 0u);
        // Every path ends within max_bounces + 1 bounces.
        for (uint32_t b = 0; b <= L.P.max_bounces; ++b) {
            const uint32_t cur = b & 1u;
            with_tree(L.tree, [&](auto t) {
                hipLaunchKernelGGL((k_wf_extend<decltype(t)::value>), grid, block, 0, stream, L.S, L.W, cur, L.spill,
                                   L.slots);
            });
            hipLaunchKernelGGL((k_wf_shade<V, T, E, D>), grid, block, 0, stream, L.S, L.P, L.W, cur, sample);
            with_tree(L.tree, [&](auto t) {
                hipLaunchKernelGGL((k_wf_shadow<V, decltype(t)::value>), grid, block, 0, stream, L.S, L.P, L.W,
                                   cur ^ 1u, L.spill, L.slots);
            });
            hipLaunchKernelGGL(k_wf_finish, grid, block, 0, stream, L.P, L.W);
        }
    }
    hipLaunchKernelGGL(k_wf_store, grid, block, 0, stream, L.P, L.W, L.accum, L.galb, L.gnorm, L.comp);
}

void launch_wavefront(const LaunchWavefront& L, const Features& f, hipStream_t stream) {
    int key = (f.volumes ? 8 : 0) | (f.textures ? 4 : 0) | (f.environment ? 2 : 0) | (f.dispersion ? 1 : 0);
    switch (key) {
    case 0: launch_wavefront_v<false, false, false, false>(L, stream); break;
    case 1: launch_wavefront_v<false, false, false, true>(L, stream); break;
    case 2: launch_wavefront_v<false, false, true, false>(L, stream); break;
    case 3: launch_wavefront_v<false, false, true, true>(L, stream); break;
    case 8: launch_wavefront_v<true, false, false, false>(L, stream); break;
    case 10: launch_wavefront_v<true, false, true, false>(L, stream); break;
    default: launch_wavefront_v<true, true, true, true>(L, stream); break;
    }
}

template <bool V, bool T, bool E, bool D>
static void launch_main_batch_v(const BatchTable& B, int tree, hipStream_t stream) {
    with_tree(tree, [&](auto t) {
        hipLaunchKernelGGL((k_main_batch<V, T, E, D, decltype(t)::value>), dim3(B.begin[B.n]), dim3(BLOCK_THREADS), 0,
                           stream, B);
    });
}

void launch_main_batch(const BatchTable& B, const Features& f, hipStream_t stream) {
    int key = (f.volumes ? 8 : 0) | (f.textures ? 4 : 0) | (f.environment ? 2 : 0) | (f.dispersion ? 1 : 0);
    switch (key) {
    case 0: launch_main_batch_v<false, false, false, false>(B, f.tree, stream); break;
    case 1: launch_main_batch_v<false, false, false, true>(B, f.tree, stream); break;
    case 2: launch_main_batch_v<false, false, true, false>(B, f.tree, stream); break;
    case 3: launch_main_batch_v<false, false, true, true>(B, f.tree, stream); break;
    case 8: launch_main_batch_v<true, false, false, false>(B, f.tree, stream); break;
    case 10: launch_main_batch_v<true, false, true, false>(B, f.tree, stream); break;
    default: launch_main_batch_v<true, true, true, true>(B, f.tree, stream); break;
    }
}

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

template <bool T, bool D>
static void launch_light_trace_v(const DevScene& S, const Params& P, const float4* emitters, uint32_t* splats,
                                 uint32_t* heads, float4* photons, uint32_t* spill, uint32_t slots, dim3 grid,
                                 int tree, hipStream_t stream) {
    with_tree(tree, [&](auto t) {
        hipLaunchKernelGGL((k_light_trace<T, D, decltype(t)::value>), grid, dim3(BLOCK_THREADS), 0, stream, S, P,
                           emitters, splats, heads, photons, spill, slots);
    });
}

void launch_light_trace(const DevScene& S, const Params& P, const float4* emitters, uint32_t* splats, uint32_t* heads,
                        float4* photons, uint32_t* spill, uint32_t slots, uint32_t count, const Features& f,
                        hipStream_t stream) {
    dim3 grid((count + BLOCK_THREADS - 1) / BLOCK_THREADS);
    if (f.textures)
        launch_light_trace_v<true, true>(S, P, emitters, splats, heads, photons, spill, slots, grid, f.tree, stream);
    else if (f.dispersion)
        launch_light_trace_v<false, true>(S, P, emitters, splats, heads, photons, spill, slots, grid, f.tree, stream);
    else
        launch_light_trace_v<false, false>(S, P, emitters, splats, heads, photons, spill, slots, grid, f.tree, stream);
}

void launch_resolve(const Params& P, uint4* splats, const float4* gathered, float4* accum, float4* comp,
                    hipStream_t stream) {
    uint32_t n = P.width * P.height;
    hipLaunchKernelGGL(k_resolve, dim3((n + BLOCK_THREADS - 1) / BLOCK_THREADS), dim3(BLOCK_THREADS), 0, stream, P,
                       splats, gathered, accum, comp);
}

void launch_fold(const Params& P, const float4* accum, const float4* galb, const float4* gnorm, const float4* comp,
                 float4* out, hipStream_t stream) {
    uint32_t n = P.width * P.height;
    hipLaunchKernelGGL(k_fold, dim3((n + BLOCK_THREADS - 1) / BLOCK_THREADS), dim3(BLOCK_THREADS), 0, stream, P,
                       accum, galb, gnorm, comp, out);
}

void launch_photon_gather(const DevScene& S, const Params& P, const uint32_t* heads, const float4* photons,
                          float4* gathered, uint32_t* spill, uint32_t slots, const Features& f, hipStream_t stream) {
    dim3 grid((P.width + 15) / 16, (P.height + 15) / 16);
    auto go = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, grid, dim3(BLOCK_THREADS), 0, stream, S, P, heads, photons, gathered, spill, slots);
    };
    with_tree(f.tree, [&](auto t) {
        constexpr int TREE = decltype(t)::value;
        if (f.textures) go(k_photon_gather<true, TREE>);
        else go(k_photon_gather<false, TREE>);
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

void launch_atrous(const Params& P, int step, const void* guides, bool fp16, const float4* in, float4* out,
                   hipStream_t stream) {
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

void launch_fill_u32(uint32_t* p, uint32_t value, size_t n, hipStream_t stream) {
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks == 0) return;
    hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)blocks), dim3(256), 0, stream, p, value, n);
}

}  // namespace rm
