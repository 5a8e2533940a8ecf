// Device code shared by the path-tracing kernels (main.hip, batch.hip,
// schedules.hip, light.hip): thread mapping, traversal stacks, the path
// integrator's steps (src/shaders.mjs:1179-1407), the per-pixel sums, and the
// launch-side helper that turns a tree kind into a template argument.
#pragma once
#include <type_traits>

#include "kernels.hpp"

namespace rm {

// ------------------------------------------------------------------ thread mapping
// A 256-thread block covers a 16×16 pixel tile as four 8×8 wave tiles, so a
// wave's primary rays are spatially coherent.
RM_D void tile_pixel_in(uint32_t t, uint32_t tile_x, uint32_t tile_y, uint32_t& x, uint32_t& y) {
    uint32_t w = t >> 6, l = t & 63u;
    x = tile_x * 16u + (w & 1u) * 8u + (l & 7u);
    y = tile_y * 16u + (w >> 1) * 8u + (l >> 3);
}
RM_D void tile_pixel_of(uint32_t t, uint32_t& x, uint32_t& y) { tile_pixel_in(t, blockIdx.x, blockIdx.y, x, y); }
RM_D void tile_pixel(uint32_t& x, uint32_t& y) { tile_pixel_of(threadIdx.x, x, y); }

// A slot for each lane with pred, from `count`, one atomic per wave; lanes
// without pred get NONE. Every lane of the wave calls it.
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

// Block shapes of the per-pixel pass (RM_OPTION_BLOCK_SHAPE): a 16×16 tile
// of four 8×8 wave squares (BLOCK_SQUARES, the default: a wave's primary
// rays stay close together), or a 64×4 tile of four 64-pixel wave rows
// (BLOCK_ROWS: a wave's per-pixel loads and stores are contiguous).
constexpr int BLOCK_SQUARES = 0, BLOCK_ROWS = 1;
template <int SHAPE>
RM_D void shaped_pixel(uint32_t& x, uint32_t& y) {
    if constexpr (SHAPE == BLOCK_ROWS) {
        x = blockIdx.x * 64u + (threadIdx.x & 63u);
        y = blockIdx.y * 4u + (threadIdx.x >> 6);
    } else {
        tile_pixel(x, y);
    }
}
template <int SHAPE>
inline dim3 shaped_grid(uint32_t width, uint32_t height) {
    if constexpr (SHAPE == BLOCK_ROWS) return dim3((width + 63) / 64, (height + 3) / 4);
    return dim3((width + 15) / 16, (height + 15) / 16);
}

// Stencil passes (the denoiser's prepare) read a neighbourhood of pixels.
// Workgroups are dealt to the 8 XCDs round-robin by linear block id, and each
// XCD has its own L2, so a plain row-major grid makes every L2 fetch nearly
// the whole image. Instead the 16×16 tiles are ordered in vertical strips
// STRIP_TILES wide (row-major inside a strip), and each XCD takes one
// contiguous run of that order: its in-flight blocks then share a compact
// footprint that fits its L2. Launch with stencil_blocks() blocks.
constexpr uint32_t STRIP_TILES = 16u;
// stencil_pixel_at takes the block's index within such a grid (a grid that
// starts on a multiple of XCDS blocks keeps each XCD's run).
RM_D bool stencil_pixel_at(uint32_t width, uint32_t height, uint32_t block, uint32_t& x, uint32_t& y) {
    const uint32_t tiles_x = (width + 15u) / 16u, tiles_y = (height + 15u) / 16u;
    const uint32_t tiles = tiles_x * tiles_y;
    const uint32_t per_xcd = (tiles + XCDS - 1u) / XCDS;
    const uint32_t t = (block % XCDS) * per_xcd + block / XCDS;
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
RM_D bool stencil_pixel(uint32_t width, uint32_t height, uint32_t& x, uint32_t& y) {
    return stencil_pixel_at(width, height, blockIdx.x, x, y);
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

// src/shaders.mjs resolve-then-gather order: pixel i's splats (fixed point,
// cleared for the next frame) and then its gathered light, when it has any,
// added to its sum a, with the error term comp[3i] when the sums are
// compensated. Returns the new sum.
RM_D float4 resolve_add(const Params& P, uint4* splats, const float4* gathered, float4* comp, uint32_t i, float4 a) {
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
    if (comp) {
        float4 e = comp[(size_t)i * 3];
        compensated_add(a, e, splat);
        if (g.w != 0.0f) compensated_add(a, e, gather);
        comp[(size_t)i * 3] = e;
    } else {
        a = make_float4(a.x + splat.x, a.y + splat.y, a.z + splat.z, a.w + splat.w);
        if (g.w != 0.0f) a = make_float4(a.x + gather.x, a.y + gather.y, a.z + gather.z, a.w + gather.w);
    }
    return a;
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

// One pixel's samples of a frame: k_main's body, shared with k_main_batch.
// LDS_DEPTH: traversal stack entries kept in LDS (lds has room for that many
// per thread of the block).
template <bool VOLUMES, bool TEXTURES, bool ENV, bool DISPERSION, int TREE, int LDS_DEPTH = MAIN_LDS_DEPTH>
RM_D void main_pixel(const DevScene& S, const Params& P, uint32_t x, uint32_t y, float4* accum, float4* guide_albedo,
                     float4* guide_normal, float4* comp, uint32_t* spill, uint32_t slots, uint2* lds) {
    if (x >= P.width || y >= P.height) return;
    const Ctx C{S, &P};
    const uint32_t index = y * P.width + x;
    auto stack = BlockStack<LDS_DEPTH, TREE>::make(lds, spill, index, slots);
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

// ------------------------------------------------------------------ path slots
// A live path's state between bounces, as the compacted, wavefront and split
// schedules keep it: SLOT_WORDS words a path, word-major over `stride` slots.
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

// Calls f with the tree kind as a compile-time constant.
template <class F>
inline void with_tree(int tree, F&& f) {
    switch (tree) {
    case TREE_WIDE: f(std::integral_constant<int, TREE_WIDE>{}); break;
    case TREE_QUANTIZED: f(std::integral_constant<int, TREE_QUANTIZED>{}); break;
    default: f(std::integral_constant<int, TREE_FP32>{}); break;
    }
}

}  // namespace rm
