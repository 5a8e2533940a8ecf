// Device-side building blocks for the Raymond kernels on gfx950: vector
// arithmetic with a fixed evaluation order, the scene records, the RNG, the
// primitive tests and the BVH traversal.
//
// Floating point follows the reference shader's formulas (src/shaders.mjs)
// operation by operation. The library is compiled with -ffp-contract=off and
// correctly rounded division and square root, so closest-hit records (t,
// primitive, barycentrics) are bit-identical to an IEEE FP32 CPU evaluation
// of the same expressions. Constants the WGSL folds at double precision
// (2π, π/4, 2π², ...) are folded the same way here.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rm {

constexpr double PI_D = 3.14159265358979;
constexpr float PI = (float)PI_D;
constexpr float TWO_PI = (float)(2.0 * PI_D);
constexpr float FOUR_PI = (float)(4.0 * PI_D);
constexpr float PI_4 = (float)(PI_D / 4.0);
constexpr float PI_2 = (float)(PI_D / 2.0);
constexpr float TWO_PI_PI = (float)(2.0 * PI_D * PI_D);
constexpr float INV_D_LINE2 = (float)(1.0 / (0.5876 * 0.5876));
constexpr uint32_t NONE = 0xffffffffu;
constexpr float T_MIN = 1e-4f;
constexpr float INF = 1e30f;
constexpr uint32_t SPHERE = 0, QUAD = 1, BOX = 2, TRIANGLE = 3;
constexpr uint32_t DIFFUSE = 0, METAL = 1, PLASTIC = 2, GLASS = 3, LIGHT = 4;
constexpr uint32_t VIEW_COLOR = 0, VIEW_ALBEDO = 1, VIEW_NORMAL = 2;
constexpr float SPLAT_SCALE = 65536.0f;
constexpr uint32_t PHOTON_TABLE = 1048576u;
constexpr uint32_t PHOTON_CAPACITY = 524288u;
constexpr uint32_t STACK_LIMIT = 64u;  // the reference's traversal stack size

#define RM_D __device__ __forceinline__

// ------------------------------------------------------------------ vectors
struct V3 { float x, y, z; };
struct V2 { float x, y; };

RM_D V3 v3(float x, float y, float z) { return {x, y, z}; }
RM_D V3 s3(float s) { return {s, s, s}; }
RM_D V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
RM_D V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
RM_D V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
RM_D V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
RM_D V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
RM_D V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
RM_D V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
RM_D V3 operator/(V3 a, V3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
RM_D float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RM_D V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
RM_D float length(V3 a) { return sqrtf(dot(a, a)); }
RM_D V3 normalize(V3 a) { float l = sqrtf(dot(a, a)); return {a.x / l, a.y / l, a.z / l}; }
RM_D V3 vmin(V3 a, V3 b) { return {fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
RM_D V3 vmax(V3 a, V3 b) { return {fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
RM_D float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
RM_D float signf(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
RM_D V3 mix(V3 a, V3 b, float t) { return a * (1.0f - t) + b * t; }
RM_D V3 reflect(V3 i, V3 n) { return i - (2.0f * dot(n, i)) * n; }
RM_D V3 refract(V3 i, V3 n, float eta) {
    float d = dot(n, i);
    float k = 1.0f - eta * eta * (1.0f - d * d);
    if (k < 0.0f) return {0.0f, 0.0f, 0.0f};
    return eta * i - (eta * d + sqrtf(k)) * n;
}
RM_D V3 xyz(float4 v) { return {v.x, v.y, v.z}; }
RM_D float luminance(V3 c) { return dot(c, v3(0.2126f, 0.7152f, 0.0722f)); }
// Saturating conversions (WGSL semantics: NaN and negatives to 0 for u32).
RM_D uint32_t to_u32(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967040.0f) return 4294967295u;
    return (uint32_t)x;
}
RM_D int32_t to_i32(float x) {
    if (x != x) return 0;
    if (x <= -2147483648.0f) return INT32_MIN;
    if (x >= 2147483520.0f) return INT32_MAX;
    return (int32_t)x;
}
RM_D uint32_t fbits(float f) { return __float_as_uint(f); }
RM_D float bitsf(uint32_t u) { return __uint_as_float(u); }
RM_D bool finite3(V3 v) {
    return (fbits(v.x) & 0x7f800000u) != 0x7f800000u && (fbits(v.y) & 0x7f800000u) != 0x7f800000u &&
           (fbits(v.z) & 0x7f800000u) != 0x7f800000u;
}

// ------------------------------------------------------------------ records
// Packed formats (64 B primitive, 32 B node, 96 B material) as the
// reference's Scene.pack() writes them (src/Scene.mjs:15-60).
struct Params {
    float cam_pos[3]; float lens_radius;
    float cam_right[3]; float tan_half_fov;
    float cam_up[3]; float focus_distance;
    float cam_forward[3]; float aspect;
    float sky_zenith[3]; float sun_omc;
    float sky_horizon[3]; uint32_t sun_on;
    float sky_ground[3]; float firefly_clamp;
    float sun_dir[3]; uint32_t light_count;
    float sun_radiance[3]; uint32_t max_bounces;
    uint32_t view, volume_count; float env_strength, env_rotation;
    uint32_t width, height, sample_offset, samples_per_frame;
    uint32_t total_samples; float exposure; uint32_t denoise, caustics;
    float photon_radius, photon_cell; uint32_t _pad5, _pad6;
};
static_assert(sizeof(Params) == 208, "the parameter block is 208 bytes");

RM_D V3 pv3(const float* p) { return {p[0], p[1], p[2]}; }

struct Mat {
    V3 color; uint32_t kind;
    V3 emission; float roughness;
    V3 checker; float checker_size;
    float ior; uint32_t texture; float tsx, tsy;
    uint32_t normal_map; float normal_scale; uint32_t emission_texture; float dispersion;
    float density, anisotropy;
};

constexpr uint32_t NO_DELTA_CUT = 0xffffffffu;

// The device-resident scene. Every pointer is device memory.
struct DevScene {
    const float4* prims;       // 4 per primitive
    const float4* nodes;       // 2 per node
    const float4* qnodes;      // 2 per node: the 8-bit child-quantized tree, or null
    const float4* wnodes;      // 4 per wide node: the 4-wide quantized tree, or null
    const float4* mats;        // 6 per material
    const float4* attr;        // vec4s
    const uint32_t* lights;    // lights, volumes, cdf bits
    const ushort4* env;        // half RGBA
    uint32_t env_w, env_h;
    const float* env_table;    // table_w × (table_h + 1)
    uint32_t table_w, table_h;
    const uchar4* textures;    // sRGB RGBA8 layers
    uint32_t tex_w, tex_h, tex_layers;
    const uchar4* nmaps;       // RGBA8 layers
    uint32_t nmap_w, nmap_h, nmap_layers;
    // The delta cut of the FP32 tree: delta_count box pairs (min, max) whose
    // union covers every leaf box holding a delta primitive (rm_scene_upload
    // builds it); NO_DELTA_CUT when there is none to rely on.
    const float4* delta_boxes = nullptr;
    uint32_t delta_count = NO_DELTA_CUT;
};

RM_D Mat load_mat(const DevScene& S, uint32_t i) {
    const float4* m = S.mats + (size_t)i * 6;
    float4 a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5];
    Mat r;
    r.color = xyz(a); r.kind = __float_as_uint(a.w);
    r.emission = xyz(b); r.roughness = b.w;
    r.checker = xyz(c); r.checker_size = c.w;
    r.ior = d.x; r.texture = __float_as_uint(d.y); r.tsx = d.z; r.tsy = d.w;
    r.normal_map = __float_as_uint(e.x); r.normal_scale = e.y; r.emission_texture = __float_as_uint(e.z);
    r.dispersion = e.w;
    r.density = f.x; r.anisotropy = f.y;
    return r;
}

// ------------------------------------------------------------------ RNG
// PCG hash, Owen-scrambled Sobol (src/shaders.mjs:131-212).
RM_D uint32_t pcg(uint32_t v) {
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
RM_D float unit(uint32_t x) { return (float)(x >> 8u) * (1.0f / 16777216.0f); }
RM_D uint32_t rbits(uint32_t x) { return __builtin_bitreverse32(x); }
RM_D uint32_t owen(uint32_t x, uint32_t seed) {
    uint32_t v = rbits(x);
    v ^= v * 0x3d20adeau;
    v += seed;
    v *= (seed >> 16u) | 1u;
    v ^= v * 0x05526c56u;
    v ^= v * 0x53a22864u;
    return rbits(v);
}
// Sobol's second dimension: the XOR of the generator columns v_k for each
// set bit k of i, where v_0 = 2^31 and v_k = v_{k-1} ^ (v_{k-1} >> 1)
// (src/shaders.mjs:160-170 walks the bits). v_k has bit 31-j set exactly when
// C(k, j) is odd, i.e. (Lucas) when the bits of j are a subset of those of k,
// so bit 31-j of the result is the XOR of i's bits over the supersets of j:
// a superset transform over the 5-bit positions, five shift-mask-XOR steps
// in place of up to 32 data-dependent iterations. Equal to the bit walk for
// every 32-bit input (checked exhaustively on the CPU).
RM_D uint32_t sobol2(uint32_t i) {
    uint32_t w = i;
    w ^= (w >> 1u) & 0x55555555u;
    w ^= (w >> 2u) & 0x33333333u;
    w ^= (w >> 4u) & 0x0F0F0F0Fu;
    w ^= (w >> 8u) & 0x00FF00FFu;
    w ^= (w >> 16u) & 0x0000FFFFu;
    return __builtin_bitreverse32(w);
}

struct Rng {
    uint32_t state;
    uint32_t seed, index, dim;
    bool qmc;

    RM_D float rand() {
        if (qmc) {
            uint32_t key = pcg(seed ^ pcg(dim));
            dim++;
            uint32_t pt = owen(index, key);
            return unit(owen(rbits(pt), pcg(key)));
        }
        state = pcg(state);
        return unit(state);
    }
    RM_D V2 rand2() {
        if (qmc) {
            uint32_t key = pcg(seed ^ pcg(dim));
            dim++;
            uint32_t pt = owen(index, key);
            return {unit(owen(rbits(pt), pcg(key))), unit(owen(sobol2(pt), pcg(key + 1u)))};
        }
        float a = rand();
        return {a, rand()};
    }
};

// ------------------------------------------------------------------ intersection
RM_D V3 safe(V3 d) {
    const float tiny = 1e-20f;
    return {fabsf(d.x) < tiny ? (d.x >= 0.0f ? tiny : -tiny) : d.x,
            fabsf(d.y) < tiny ? (d.y >= 0.0f ? tiny : -tiny) : d.y,
            fabsf(d.z) < tiny ? (d.z >= 0.0f ? tiny : -tiny) : d.z};
}

// (t or -1, u, v) for primitive i between T_MIN and t_max (src/shaders.mjs:262-359).
RM_D V3 intersect(const DevScene& S, uint32_t i, V3 o, V3 d, float t_max) {
    const float4* p = S.prims + (size_t)i * 4;
    const V3 miss = {-1.0f, 0.0f, 0.0f};
    uint32_t kind = __float_as_uint(p[0].x);
    float4 p0 = p[1];
    if (kind == SPHERE) {
        V3 oc = o - xyz(p0);
        float r = p0.w;
        float a = dot(d, d);
        float half_b = dot(oc, d);
        float c = dot(oc, oc) - r * r;
        float disc = half_b * half_b - a * c;
        if (disc < 0.0f) return miss;
        float sq = sqrtf(disc);
        float t = (-half_b - sq) / a;
        if (t <= T_MIN || t >= t_max) {
            t = (-half_b + sq) / a;
            if (t <= T_MIN || t >= t_max) return miss;
        }
        return {t, 0.0f, 0.0f};
    }
    float4 p1 = p[2], p2 = p[3];
    if (kind == TRIANGLE) {
        V3 v0 = xyz(p0);
        V3 e1 = xyz(p1) - v0, e2 = xyz(p2) - v0;
        V3 pv = cross(d, e2);
        float det = dot(e1, pv);
        if (fabsf(det) < 1e-12f) return miss;
        float inv_det = 1.0f / det;
        V3 tv = o - v0;
        float bu = dot(tv, pv) * inv_det;
        if (bu < 0.0f || bu > 1.0f) return miss;
        V3 qv = cross(tv, e1);
        float bv = dot(d, qv) * inv_det;
        if (bv < 0.0f || bu + bv > 1.0f) return miss;
        float t = dot(e2, qv) * inv_det;
        if (t <= T_MIN || t >= t_max) return miss;
        return {t, bu, bv};
    }
    if (kind == QUAD) {
        V3 q = xyz(p0), qu = xyz(p1), qv = xyz(p2);
        V3 n = cross(qu, qv);
        float denom = dot(n, d);
        if (fabsf(denom) < 1e-12f) return miss;
        float t = dot(n, q - o) / denom;
        if (t <= T_MIN || t >= t_max) return miss;
        V3 h = o + d * t - q;
        float nn = dot(n, n);
        float alpha = dot(n, cross(h, qv)) / nn;
        float beta = dot(n, cross(qu, h)) / nn;
        if (alpha < 0.0f || alpha > 1.0f || beta < 0.0f || beta > 1.0f) return miss;
        return {t, alpha, beta};
    }
    if (kind == BOX) {
        V3 lo = {dot(xyz(p0), o) + p0.w, dot(xyz(p1), o) + p1.w, dot(xyz(p2), o) + p2.w};
        V3 ld = safe({dot(xyz(p0), d), dot(xyz(p1), d), dot(xyz(p2), d)});
        V3 inv = {1.0f / ld.x, 1.0f / ld.y, 1.0f / ld.z};
        V3 t0 = (s3(-1.0f) - lo) * inv;
        V3 t1 = (s3(1.0f) - lo) * inv;
        V3 tmin = vmin(t0, t1), tmax = vmax(t0, t1);
        float t_near = fmaxf(fmaxf(tmin.x, tmin.y), tmin.z);
        float t_far = fminf(fminf(tmax.x, tmax.y), tmax.z);
        if (t_near > t_far) return miss;
        float t = t_near > T_MIN ? t_near : t_far;
        if (t <= T_MIN || t >= t_max) return miss;
        return {t, 0.0f, 0.0f};
    }
    return miss;
}

// Entry distance into a node's box, clamped to [0, t_max], else INF.
RM_D float slabs(float4 lo, float4 hi, V3 o, V3 inv, float t_max) {
    V3 t0 = (xyz(lo) - o) * inv;
    V3 t1 = (xyz(hi) - o) * inv;
    V3 tmin = vmin(t0, t1), tmax = vmax(t0, t1);
    float t_near = fmaxf(fmaxf(tmin.x, tmin.y), fmaxf(tmin.z, 0.0f));
    float t_far = fminf(fminf(tmax.x, tmax.y), fminf(tmax.z, t_max));
    return t_near <= t_far ? t_near : INF;
}

struct Hit { float t; uint32_t prim; float u, v; };

// A lane's traversal stack: the first LDS_DEPTH entries in LDS, laid out
// entry-major across the block ([entry][thread]) so a wave's lanes hit
// distinct banks; deeper entries spill to this thread's slice of global
// memory. An entry is a pushed node's (first, count), which the parent's
// child test has already loaded, so a pop costs no global load. Capacity is
// the reference's 64 entries; a full stack drops the far child exactly as the
// reference does.
// The trees trace() can walk.
constexpr int TREE_FP32 = 0;        // the packed FP32 nodes
constexpr int TREE_QUANTIZED = 1;   // the 8-bit child-quantized binary records
constexpr int TREE_WIDE = 2;        // the 4-wide 8-bit quantized records

// TREE selects the tree trace() walks; an entry is what that walk pushes.
template <int LDS_DEPTH, int BLOCK, int TREE = TREE_FP32>
struct Stack {
    static constexpr int tree = TREE;
    static constexpr int lds_depth = LDS_DEPTH;
    static constexpr int block = BLOCK;
    uint2* lds;      // &shared[0][threadIdx.x]
    uint2* spill;    // this thread's column of an [entry][slot] spill array
    uint32_t stride; // slots in the spill array
    // The LDS entries go through an LDS-space pointer and the spilled ones
    // through a volatile one: otherwise the compiler merges the two branches'
    // stores (and loads) into one flat access, which waits on both the
    // vector-memory and the LDS counters, so that every traversal step waited
    // for all loads in flight.
    using Local = __attribute__((address_space(3))) uint64_t;
    RM_D Local* local() const { return (Local*)lds; }
    RM_D void put(uint32_t sp, uint2 v) {
        if (LDS_DEPTH >= (int)STACK_LIMIT || sp < (uint32_t)LDS_DEPTH) {
            local()[sp * BLOCK] = (uint64_t)v.y << 32 | v.x;
        } else {
            volatile uint2* at = spill + (size_t)(sp - LDS_DEPTH) * stride;
            at->x = v.x;
            at->y = v.y;
        }
    }
    RM_D uint2 get(uint32_t sp) const {
        if (LDS_DEPTH >= (int)STACK_LIMIT || sp < (uint32_t)LDS_DEPTH) {
            const uint64_t e = local()[sp * BLOCK];
            return make_uint2((uint32_t)e, (uint32_t)(e >> 32));
        }
        const volatile uint2* at = spill + (size_t)(sp - LDS_DEPTH) * stride;
        return make_uint2(at->x, at->y);
    }
};

// Traversal counters, compiled in only when asked for.
struct NoCount { RM_D void nodes(uint32_t) {} RM_D void prims(uint32_t) {} };
struct Count {
    uint32_t n = 0, p = 0;
    RM_D void nodes(uint32_t k) { n += k; }
    RM_D void prims(uint32_t k) { p += k; }
};

// The loop the three walks share. A lane holds a leaf when count > 0 (its
// primitives [first, first + count)) and an inner node otherwise; node(t)
// takes one inner node, moving to the nearer child entered before t and
// pushing the rest, and returns false when the ray enters none; pop() takes
// the top entry. Closest-hit rays take inner nodes in a loop of their own
// until the lane holds a leaf (or is done), then the leaf, so the lanes of a
// wave test boxes together and primitives together: 10% faster on triangle
// meshes and glass than one node or leaf per pass, and within 1% on small
// sphere scenes. Shadow rays, which stop at their first hit, are faster with
// one node or leaf per pass. Either way a ray visits its nodes and
// primitives in the same order.
//
// TIES (closest hit over a quantized tree): a primitive hit at exactly the
// nearest t so far sets `tie` (a nearer hit clears it). The quantized boxes
// can order two overlapping children differently from the FP32 boxes, and
// an equal-t hit goes to whichever primitive is reached first, so a ray that
// ends with a tie is resolved again by the FP32 walk. To see such hits, a
// primitive is tested against the float just above the nearest t; that
// bound changes which hits are accepted, never the t, u, v computed.
template <bool ANY, bool TIES = false, class NODE, class POP, class COUNT>
RM_D Hit walk(const DevScene& S, V3 o, V3 d, float t_max, uint32_t& first, uint32_t& count, uint32_t& sp,
              NODE&& node, POP&& pop, COUNT& cnt, bool* tie = nullptr) {
    Hit hit = {t_max, NONE, 0.0f, 0.0f};
    if constexpr (ANY) {
        for (;;) {
            if (count > 0u) {
                for (uint32_t i = first; i < first + count; i++) {
                    cnt.prims(1);
                    V3 h = intersect(S, i, o, d, hit.t);
                    if (h.x > 0.0f) return {h.x, i, h.y, h.z};
                }
            } else if (node(hit.t)) {
                continue;
            }
            if (sp == 0u) break;
            pop();
        }
    } else {
        for (;;) {
            bool done = false;
            while (count == 0u) {
                if (node(hit.t)) continue;
                if (sp == 0u) {
                    done = true;
                    break;
                }
                pop();
            }
            if (done) break;
            for (uint32_t i = first; i < first + count; i++) {
                cnt.prims(1);
                if constexpr (TIES) {
                    // Positive and finite: the float just above is one ulp up.
                    const float bound = hit.prim == NONE ? hit.t : __uint_as_float(__float_as_uint(hit.t) + 1u);
                    V3 h = intersect(S, i, o, d, bound);
                    if (h.x > 0.0f) {
                        if (h.x < hit.t || hit.prim == NONE) {
                            hit = {h.x, i, h.y, h.z};
                            *tie = false;
                        } else {
                            *tie = true;
                        }
                    }
                } else {
                    V3 h = intersect(S, i, o, d, hit.t);
                    if (h.x > 0.0f) hit = {h.x, i, h.y, h.z};
                }
            }
            if (sp == 0u) break;
            pop();
        }
    }
    return hit;
}

// The reference's trace() (src/shaders.mjs:389-447): nearest hit before
// t_max, or with ANY, the first hit found. Visits the nearer child first
// (ties to the left), pushes the farther when the ray enters both and the
// stack has room, and never re-tests a popped node, so every hit record and
// every tie resolves as the reference's does. The current node's (first,
// count) come along from the child test, so a node record is loaded once.
template <bool ANY, class STACK, class COUNT>
RM_D Hit trace_fp32(const DevScene& S, V3 o, V3 d, float t_max, STACK& stack, COUNT& cnt) {
    V3 sd = safe(d);
    V3 inv = {1.0f / sd.x, 1.0f / sd.y, 1.0f / sd.z};
    const float4* N = S.nodes;
    cnt.nodes(1);
    float4 r0 = N[0], r1 = N[1];
    if (slabs(r0, r1, o, inv, t_max) >= INF) return {t_max, NONE, 0.0f, 0.0f};
    uint32_t first = __float_as_uint(r0.w), count = __float_as_uint(r1.w);
    uint32_t sp = 0;
    auto node = [&](float t) {
        cnt.nodes(2);
        const float4* c = N + (size_t)first * 2;
        float4 l0 = c[0], l1 = c[1], q0 = c[2], q1 = c[3];
        float t_left = slabs(l0, l1, o, inv, t);
        float t_right = slabs(q0, q1, o, inv, t);
        bool left_first = t_left <= t_right;
        if (!(fminf(t_left, t_right) < INF)) return false;
        if (fmaxf(t_left, t_right) < INF && sp < STACK_LIMIT) {
            stack.put(sp, make_uint2(__float_as_uint(left_first ? q0.w : l0.w),
                                     __float_as_uint(left_first ? q1.w : l1.w)));
            sp++;
        }
        first = __float_as_uint(left_first ? l0.w : q0.w);
        count = __float_as_uint(left_first ? l1.w : q1.w);
        return true;
    };
    auto pop = [&]() {
        sp--;
        uint2 e = stack.get(sp);
        first = e.x;
        count = e.y;
    };
    return walk<ANY>(S, o, d, t_max, first, count, sp, node, pop, cnt);
}

// A quantized tree's walk. Any-hit rays need no care: the decoded boxes
// enclose the FP32 ones, so every primitive the FP32 walk could report is
// reachable and occlusion agrees. A closest-hit ray that ends on an exact-t
// tie is traced again over the FP32 nodes (resident beside the quantized
// ones), which picks the primitive the FP32 walk picks; the walk is done,
// so its stack is free to reuse.
template <bool ANY, class NODE, class POP, class STACK, class COUNT>
RM_D Hit walk_resolving_ties(const DevScene& S, V3 o, V3 d, float t_max, uint32_t& first, uint32_t& count,
                             uint32_t& sp, NODE&& node, POP&& pop, STACK& stack, COUNT& cnt) {
    if constexpr (ANY) {
        return walk<true>(S, o, d, t_max, first, count, sp, node, pop, cnt);
    } else {
        bool tie = false;
        Hit h = walk<false, true>(S, o, d, t_max, first, count, sp, node, pop, cnt, &tie);
        if (!tie) return h;
        Stack<STACK::lds_depth, STACK::block, TREE_FP32> flat{stack.lds, stack.spill, stack.stride};
        return trace_fp32<false>(S, o, d, t_max, flat, cnt);
    }
}

// One axis of a quantized child box: origin + q·step, q·step exact (an
// 8-bit integer times a power of two), so only the sum rounds — the
// expression the host rounded outward against.
RM_D float dequant(float origin, uint32_t q, float step) { return origin + (float)q * step; }

RM_D float step_of(uint32_t biased) { return __uint_as_float(biased << 23); }

// trace() over the 8-bit child-quantized tree (rm_bvh_quantize): the same
// walk — nearer child first, ties to the left, the farther pushed when both
// are entered and the stack has room, no re-test on pop — with each inner
// node's two child boxes decoded from one 32-byte record. Decoded boxes
// enclose the FP32 ones, so the leaves the FP32 walk tests are all tested
// here and the nearest hit is the same; only the visiting order of boxes
// that overlap along the ray can differ, which matters for hits at exactly
// equal t.
template <bool ANY, class STACK, class COUNT>
RM_D Hit trace_q(const DevScene& S, V3 o, V3 d, float t_max, STACK& stack, COUNT& cnt) {
    V3 sd = safe(d);
    V3 inv = {1.0f / sd.x, 1.0f / sd.y, 1.0f / sd.z};
    const float4* Q = S.qnodes;
    cnt.nodes(1);
    float4 r0 = S.nodes[0], r1 = S.nodes[1];
    if (slabs(r0, r1, o, inv, t_max) >= INF) return {t_max, NONE, 0.0f, 0.0f};
    uint32_t at = 0u, first = __float_as_uint(r0.w), count = __float_as_uint(r1.w);
    uint32_t sp = 0;
    // Moves to record `next`, loading a leaf's range from it.
    auto go = [&](uint32_t next, uint32_t leaf) {
        at = next;
        count = 0u;
        if (leaf) {
            const float4 e = Q[(size_t)next * 2];
            first = __float_as_uint(e.x);
            count = __float_as_uint(e.y);
        }
    };
    auto node = [&](float t) {
        cnt.nodes(2);
        const float4 a = Q[(size_t)at * 2], b = Q[(size_t)at * 2 + 1];
        const uint32_t meta = __float_as_uint(a.w);
        const float sx = step_of(meta & 0xffu), sy = step_of(meta >> 8 & 0xffu), sz = step_of(meta >> 16 & 0xffu);
        const uint32_t w0 = __float_as_uint(b.x), w1 = __float_as_uint(b.y), w2 = __float_as_uint(b.z);
        const float4 l0 = make_float4(dequant(a.x, w0 & 0xffu, sx), dequant(a.y, w0 >> 8 & 0xffu, sy),
                                      dequant(a.z, w0 >> 16 & 0xffu, sz), 0.0f);
        const float4 l1 = make_float4(dequant(a.x, w0 >> 24, sx), dequant(a.y, w1 & 0xffu, sy),
                                      dequant(a.z, w1 >> 8 & 0xffu, sz), 0.0f);
        const float4 q0 = make_float4(dequant(a.x, w1 >> 16 & 0xffu, sx), dequant(a.y, w1 >> 24, sy),
                                      dequant(a.z, w2 & 0xffu, sz), 0.0f);
        const float4 q1 = make_float4(dequant(a.x, w2 >> 8 & 0xffu, sx), dequant(a.y, w2 >> 16 & 0xffu, sy),
                                      dequant(a.z, w2 >> 24, sz), 0.0f);
        float t_left = slabs(l0, l1, o, inv, t);
        float t_right = slabs(q0, q1, o, inv, t);
        bool left_first = t_left <= t_right;
        if (!(fminf(t_left, t_right) < INF)) return false;
        const uint32_t c = __float_as_uint(b.w);
        if (fmaxf(t_left, t_right) < INF && sp < STACK_LIMIT) {
            stack.put(sp, make_uint2(left_first ? c + 1u : c, meta >> (left_first ? 25 : 24) & 1u));
            sp++;
        }
        go(left_first ? c : c + 1u, meta >> (left_first ? 24 : 25) & 1u);
        return true;
    };
    auto pop = [&]() {
        sp--;
        const uint2 e = stack.get(sp);
        go(e.x, e.y);
    };
    return walk_resolving_ties<ANY>(S, o, d, t_max, first, count, sp, node, pop, stack, cnt);
}

// Byte i (a constant once unrolled) of six packed words.
RM_D uint32_t byte_of(const uint32_t (&w)[6], int i) { return w[i >> 2] >> (8 * (i & 3)) & 0xffu; }

// trace() over the 4-wide quantized tree (rm_bvh_quantize_wide): each step
// decodes two to four child boxes from one 64-byte record, visits the
// nearest child entered and pushes the others, nearer ones on top. Leaf
// ranges sit in the parent's record, so a leaf costs no load of its own.
// Decoded boxes enclose the FP32 ones, so every leaf the FP32 walk reaches
// is reached and the nearest t is the same; the order differs, which matters
// only for hits at exactly equal t. The producer bounds the pushes on any
// path by the 64-entry stack, so none is ever dropped.
template <bool ANY, class STACK, class COUNT>
RM_D Hit trace_w(const DevScene& S, V3 o, V3 d, float t_max, STACK& stack, COUNT& cnt) {
    V3 sd = safe(d);
    V3 inv = {1.0f / sd.x, 1.0f / sd.y, 1.0f / sd.z};
    cnt.nodes(1);
    float4 r0 = S.nodes[0], r1 = S.nodes[1];
    if (slabs(r0, r1, o, inv, t_max) >= INF) return {t_max, NONE, 0.0f, 0.0f};
    // count > 0: a leaf's primitives [first, first + count); 0: wide node `first`.
    uint32_t first = __float_as_uint(r1.w) > 0u ? __float_as_uint(r0.w) : 0u, count = __float_as_uint(r1.w);
    uint32_t sp = 0;
    auto node = [&](float tm) {
        const float4* W = S.wnodes + (size_t)first * 4;
        const float4 a = W[0], b = W[1], c = W[2], e = W[3];
        const uint32_t meta = __float_as_uint(a.w);
        const uint32_t n = meta >> 28;
        cnt.nodes(n);
        const float sx = step_of(meta & 0xffu), sy = step_of(meta >> 8 & 0xffu), sz = step_of(meta >> 16 & 0xffu);
        const uint32_t bw[6] = {__float_as_uint(b.x), __float_as_uint(b.y), __float_as_uint(b.z),
                                __float_as_uint(b.w), __float_as_uint(c.x), __float_as_uint(c.y)};
        const uint32_t ref[4] = {__float_as_uint(c.z), __float_as_uint(c.w), __float_as_uint(e.x),
                                 __float_as_uint(e.y)};
        const uint32_t counts = __float_as_uint(e.z);
        float t[4];
        uint32_t id[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            id[k] = (uint32_t)k;
            t[k] = INF;
            if ((uint32_t)k < n) {
                const float4 lo = make_float4(dequant(a.x, byte_of(bw, 6 * k), sx),
                                              dequant(a.y, byte_of(bw, 6 * k + 1), sy),
                                              dequant(a.z, byte_of(bw, 6 * k + 2), sz), 0.0f);
                const float4 hi = make_float4(dequant(a.x, byte_of(bw, 6 * k + 3), sx),
                                              dequant(a.y, byte_of(bw, 6 * k + 4), sy),
                                              dequant(a.z, byte_of(bw, 6 * k + 5), sz), 0.0f);
                t[k] = slabs(lo, hi, o, inv, tm);
            }
        }
        // Nearest first; equal distances in slot order.
        auto order = [&](int i, int j) {
            if (t[j] < t[i] || (t[j] == t[i] && id[j] < id[i])) {
                const float tt = t[i]; t[i] = t[j]; t[j] = tt;
                const uint32_t ii = id[i]; id[i] = id[j]; id[j] = ii;
            }
        };
        order(0, 1); order(2, 3); order(0, 2); order(1, 3); order(1, 2);
        if (!(t[0] < INF)) return false;
#pragma unroll
        for (int k = 3; k >= 1; --k) {
            if (t[k] < INF && sp < STACK_LIMIT) {
                const uint32_t leaf = meta >> (24 + id[k]) & 1u;
                stack.put(sp, make_uint2(ref[id[k]], leaf ? counts >> (8 * id[k]) & 0xffu : 0u));
                sp++;
            }
        }
        const uint32_t leaf = meta >> (24 + id[0]) & 1u;
        first = ref[id[0]];
        count = leaf ? counts >> (8 * id[0]) & 0xffu : 0u;
        return true;
    };
    auto pop = [&]() {
        sp--;
        const uint2 en = stack.get(sp);
        first = en.x;
        count = en.y;
    };
    return walk_resolving_ties<ANY>(S, o, d, t_max, first, count, sp, node, pop, stack, cnt);
}

// The walk the stack's type asks for.
template <bool ANY, class STACK, class COUNT>
RM_D Hit trace(const DevScene& S, V3 o, V3 d, float t_max, STACK& stack, COUNT& cnt) {
    if constexpr (STACK::tree == TREE_WIDE)
        return trace_w<ANY>(S, o, d, t_max, stack, cnt);
    else if constexpr (STACK::tree == TREE_QUANTIZED)
        return trace_q<ANY>(S, o, d, t_max, stack, cnt);
    else
        return trace_fp32<ANY>(S, o, d, t_max, stack, cnt);
}

}  // namespace rm
