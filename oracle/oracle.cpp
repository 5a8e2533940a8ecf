// CPU oracle for the Raymond rendering path — TEST INFRASTRUCTURE ONLY (see
// oracle.h). Each function cites the reference WGSL it restates; FP32 with
// the shader's operation order and WGSL's const-expression folding (PI-based
// constants are folded in double, then rounded once to f32).
//
// Build: g++ -O2 -std=c++17 -ffp-contract=off -fno-fast-math -shared -fPIC -pthread
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------- constants
// src/shaders.mjs:44-47
constexpr double PI_D = 3.14159265358979;
const float PI = (float)PI_D;
const float TWO_PI = (float)(2.0 * PI_D);
const float FOUR_PI = (float)(4.0 * PI_D);
const float PI_4 = (float)(PI_D / 4.0);
const float PI_2 = (float)(PI_D / 2.0);
const float TWO_PI_PI = (float)(2.0 * PI_D * PI_D);
constexpr uint32_t NONE = 0xffffffffu;
constexpr float T_MIN = 1e-4f;
constexpr float INF = 1e30f;
constexpr uint32_t SPHERE = 0, QUAD = 1, BOX = 2, TRIANGLE = 3;
constexpr uint32_t DIFFUSE = 0, METAL = 1, PLASTIC = 2, GLASS = 3, LIGHT = 4;
constexpr uint32_t VIEW_COLOR = 0, VIEW_ALBEDO = 1, VIEW_NORMAL = 2;
constexpr float SPLAT_SCALE = 65536.0f;
constexpr uint32_t PHOTON_TABLE = 1048576u;
constexpr uint32_t PHOTON_CAPACITY = 524288u;

// ---------------------------------------------------------------- vectors
struct V3 { float x, y, z; };
struct V2 { float x, y; };
struct V4 { float x, y, z, w; };

inline V3 v3(float x, float y, float z) { return {x, y, z}; }
inline V3 splat3(float s) { return {s, s, s}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3 operator/(V3 a, V3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
// dot = (x·x' + y·y') + z·z', the fixed order both sides use.
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float length(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalize(V3 a) { float l = std::sqrt(dot(a, a)); return {a.x / l, a.y / l, a.z / l}; }
inline float fmin_(float a, float b) { return std::fmin(a, b); }
inline float fmax_(float a, float b) { return std::fmax(a, b); }
inline V3 vmin(V3 a, V3 b) { return {fmin_(a.x, b.x), fmin_(a.y, b.y), fmin_(a.z, b.z)}; }
inline V3 vmax(V3 a, V3 b) { return {fmax_(a.x, b.x), fmax_(a.y, b.y), fmax_(a.z, b.z)}; }
inline float clampf(float x, float lo, float hi) { return fmin_(fmax_(x, lo), hi); }
inline float signf(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
inline V3 mix(V3 a, V3 b, float t) { return a * (1.0f - t) + b * t; }
inline V3 reflect(V3 i, V3 n) { return i - (2.0f * dot(n, i)) * n; }
// WGSL refract(e1, e2, e3): zero when totally reflected.
inline V3 refract(V3 i, V3 n, float eta) {
    float d = dot(n, i);
    float k = 1.0f - eta * eta * (1.0f - d * d);
    if (k < 0.0f) return {0.0f, 0.0f, 0.0f};
    return eta * i - (eta * d + std::sqrt(k)) * n;
}
// Saturating float → integer conversions, as WGSL specifies.
inline uint32_t to_u32(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967040.0f) return 4294967295u;
    return (uint32_t)x;
}
inline int32_t to_i32(float x) {
    if (x != x) return 0;
    if (x <= -2147483648.0f) return INT32_MIN;
    if (x >= 2147483520.0f) return INT32_MAX;
    return (int32_t)x;
}
inline float bits_f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
inline uint32_t f_bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float luminance(V3 c) { return dot(c, v3(0.2126f, 0.7152f, 0.0722f)); }
// finite(): no component with an all-ones exponent (src/shaders.mjs:1833-1836).
inline bool finite3(V3 v) {
    return (f_bits(v.x) & 0x7f800000u) != 0x7f800000u && (f_bits(v.y) & 0x7f800000u) != 0x7f800000u &&
           (f_bits(v.z) & 0x7f800000u) != 0x7f800000u;
}

float half_to_float(uint16_t h) {
    uint32_t s = (h >> 15) & 1, e = (h >> 10) & 31, m = h & 1023;
    float v;
    if (e == 0) v = std::ldexp((float)m, -24);
    else if (e == 31) v = m ? NAN : INFINITY;
    else v = std::ldexp((float)(m | 1024), (int)e - 25);
    return s ? -v : v;
}

// ---------------------------------------------------------------- records
struct Prim { uint32_t kind, material, light, attributes; float p0[4], p1[4], p2[4]; };
struct Node { float mn[3]; uint32_t first; float mx[3]; uint32_t count; };
struct Mat {
    float color[3]; uint32_t kind;
    float emission[3]; float roughness;
    float checker[3]; float checker_size;
    float ior; uint32_t texture; float texture_scale[2];
    uint32_t normal_map; float normal_scale; uint32_t emission_texture; float dispersion;
    float density; float anisotropy; float _m0, _m1;
};
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
static_assert(sizeof(Prim) == 64, "primitive is 64 bytes");
static_assert(sizeof(Node) == 32, "node is 32 bytes");
static_assert(sizeof(Mat) == 96, "material is 96 bytes");
static_assert(sizeof(Params) == 208, "params is 208 bytes");

inline V3 xyz(const float* p) { return {p[0], p[1], p[2]}; }

struct Scene {
    const Prim* prims;
    const Node* nodes;
    const Mat* mats;
    const float* attr;
    const uint32_t* lights;
    const OracleScene* raw;
};

Scene view(const OracleScene* s) {
    return {reinterpret_cast<const Prim*>(s->prims), reinterpret_cast<const Node*>(s->nodes),
            reinterpret_cast<const Mat*>(s->materials), s->attributes, s->lights, s};
}

// ---------------------------------------------------------------- RNG
// src/shaders.mjs:131-212
inline uint32_t pcg(uint32_t v) {
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
inline float unit(uint32_t x) { return (float)(x >> 8u) * (1.0f / 16777216.0f); }
inline uint32_t reverse_bits(uint32_t x) {
    x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
    x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
    x = ((x >> 4) & 0x0f0f0f0fu) | ((x & 0x0f0f0f0fu) << 4);
    x = ((x >> 8) & 0x00ff00ffu) | ((x & 0x00ff00ffu) << 8);
    return (x >> 16) | (x << 16);
}
inline uint32_t owen(uint32_t x, uint32_t seed) {
    uint32_t v = reverse_bits(x);
    v ^= v * 0x3d20adeau;
    v += seed;
    v *= (seed >> 16u) | 1u;
    v ^= v * 0x05526c56u;
    v ^= v * 0x53a22864u;
    return reverse_bits(v);
}
inline uint32_t sobol2(uint32_t i) {
    uint32_t r = 0u, v = 0x80000000u;
    for (uint32_t j = i; j != 0u; j >>= 1u) {
        if (j & 1u) r ^= v;
        v ^= v >> 1u;
    }
    return r;
}

struct Rng {
    uint32_t state = 0;
    bool qmc = false;
    uint32_t seed = 0, index = 0, dim = 0;

    void next(uint32_t& key, uint32_t& point) {
        key = pcg(seed ^ pcg(dim));
        dim++;
        point = owen(index, key);
    }
    float rand() {
        if (qmc) {
            uint32_t k, p;
            next(k, p);
            return unit(owen(reverse_bits(p), pcg(k)));
        }
        state = pcg(state);
        return unit(state);
    }
    V2 rand2() {
        if (qmc) {
            uint32_t k, p;
            next(k, p);
            return {unit(owen(reverse_bits(p), pcg(k))), unit(owen(sobol2(p), pcg(k + 1u)))};
        }
        float a = rand();
        return {a, rand()};
    }
};

// ---------------------------------------------------------------- samplers
// src/shaders.mjs:217-256
struct Frame { V3 c0, c1, c2; };
inline Frame frame(V3 n) {
    float s = n.z >= 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (s + n.z);
    float b = n.x * n.y * a;
    return {{1.0f + s * n.x * n.x * a, s * b, -s * n.x}, {b, s + n.y * n.y * a, -n.y}, n};
}
inline V3 fmul(const Frame& f, V3 l) { return f.c0 * l.x + f.c1 * l.y + f.c2 * l.z; }

V3 cosine_hemisphere(Rng& g, V3 n) {
    V2 u = g.rand2();
    float r = std::sqrt(u.x);
    float phi = TWO_PI * u.y;
    V3 local = {r * std::cos(phi), r * std::sin(phi), std::sqrt(fmax_(0.0f, 1.0f - r * r))};
    return normalize(fmul(frame(n), local));
}

V3 sample_cone(Rng& g, V3 axis, float omc) {
    V2 u = g.rand2();
    float cos_t = 1.0f - u.x * omc;
    float sin_t = std::sqrt(fmax_(0.0f, 1.0f - cos_t * cos_t));
    float phi = TWO_PI * u.y;
    return normalize(fmul(frame(axis), {sin_t * std::cos(phi), sin_t * std::sin(phi), cos_t}));
}

V2 concentric_disk(Rng& g) {
    V2 r = g.rand2();
    V2 u = {2.0f * r.x - 1.0f, 2.0f * r.y - 1.0f};
    if (u.x == 0.0f && u.y == 0.0f) return {0.0f, 0.0f};
    if (std::fabs(u.x) > std::fabs(u.y)) {
        float phi = PI_4 * (u.y / u.x);
        return {u.x * std::cos(phi), u.x * std::sin(phi)};
    }
    float phi = PI_2 - PI_4 * (u.x / u.y);
    return {u.y * std::cos(phi), u.y * std::sin(phi)};
}

// ---------------------------------------------------------------- intersection
// src/shaders.mjs:262-378
inline V3 safe(V3 d) {
    const float tiny = 1e-20f;
    auto one = [&](float c) { return std::fabs(c) < tiny ? (c >= 0.0f ? tiny : -tiny) : c; };
    return {one(d.x), one(d.y), one(d.z)};
}

struct Counters { uint64_t nodes = 0, prims = 0; };

V3 intersect(const Scene& S, uint32_t i, V3 o, V3 d, float t_max) {
    const Prim& p = S.prims[i];
    const V3 miss = {-1.0f, 0.0f, 0.0f};
    switch (p.kind) {
    case SPHERE: {
        V3 oc = o - xyz(p.p0);
        float r = p.p0[3];
        float a = dot(d, d);
        float half_b = dot(oc, d);
        float c = dot(oc, oc) - r * r;
        float disc = half_b * half_b - a * c;
        if (disc < 0.0f) return miss;
        float sq = std::sqrt(disc);
        float t = (-half_b - sq) / a;
        if (t <= T_MIN || t >= t_max) {
            t = (-half_b + sq) / a;
            if (t <= T_MIN || t >= t_max) return miss;
        }
        return {t, 0.0f, 0.0f};
    }
    case QUAD: {
        V3 q = xyz(p.p0), qu = xyz(p.p1), qv = xyz(p.p2);
        V3 n = cross(qu, qv);
        float denom = dot(n, d);
        if (std::fabs(denom) < 1e-12f) return miss;
        float t = dot(n, q - o) / denom;
        if (t <= T_MIN || t >= t_max) return miss;
        V3 h = o + d * t - q;
        float nn = dot(n, n);
        float alpha = dot(n, cross(h, qv)) / nn;
        float beta = dot(n, cross(qu, h)) / nn;
        if (alpha < 0.0f || alpha > 1.0f || beta < 0.0f || beta > 1.0f) return miss;
        return {t, alpha, beta};
    }
    case BOX: {
        V3 lo = {dot(xyz(p.p0), o) + p.p0[3], dot(xyz(p.p1), o) + p.p1[3], dot(xyz(p.p2), o) + p.p2[3]};
        V3 ld = safe({dot(xyz(p.p0), d), dot(xyz(p.p1), d), dot(xyz(p.p2), d)});
        V3 inv = {1.0f / ld.x, 1.0f / ld.y, 1.0f / ld.z};
        V3 t0 = (splat3(-1.0f) - lo) * inv;
        V3 t1 = (splat3(1.0f) - lo) * inv;
        V3 tmin = vmin(t0, t1), tmax = vmax(t0, t1);
        float t_near = fmax_(fmax_(tmin.x, tmin.y), tmin.z);
        float t_far = fmin_(fmin_(tmax.x, tmax.y), tmax.z);
        if (t_near > t_far) return miss;
        float t = t_near > T_MIN ? t_near : t_far;
        if (t <= T_MIN || t >= t_max) return miss;
        return {t, 0.0f, 0.0f};
    }
    case TRIANGLE: {
        V3 v0 = xyz(p.p0);
        V3 e1 = xyz(p.p1) - v0, e2 = xyz(p.p2) - v0;
        V3 pv = cross(d, e2);
        float det = dot(e1, pv);
        if (std::fabs(det) < 1e-12f) return miss;
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
    default:
        return miss;
    }
}

inline float slabs(const Node& n, V3 o, V3 inv, float t_max) {
    V3 t0 = (xyz(n.mn) - o) * inv;
    V3 t1 = (xyz(n.mx) - o) * inv;
    V3 tmin = vmin(t0, t1), tmax = vmax(t0, t1);
    float t_near = fmax_(fmax_(tmin.x, tmin.y), fmax_(tmin.z, 0.0f));
    float t_far = fmin_(fmin_(tmax.x, tmax.y), fmin_(tmax.z, t_max));
    return t_near <= t_far ? t_near : INF;
}

struct Hit { float t; uint32_t prim; float u, v; };

// src/shaders.mjs:389-447
Hit trace(const Scene& S, V3 o, V3 d, float t_max, bool any_hit, Counters* cnt = nullptr) {
    Hit hit = {t_max, NONE, 0.0f, 0.0f};
    V3 sd = safe(d);
    V3 inv = {1.0f / sd.x, 1.0f / sd.y, 1.0f / sd.z};
    if (cnt) cnt->nodes++;
    if (slabs(S.nodes[0], o, inv, t_max) >= INF) return hit;
    uint32_t stack[64];
    uint32_t sp = 0, current = 0;
    for (;;) {
        const Node& n = S.nodes[current];
        bool descend = false;
        if (n.count > 0u) {
            for (uint32_t i = n.first; i < n.first + n.count; i++) {
                if (cnt) cnt->prims++;
                V3 h = intersect(S, i, o, d, hit.t);
                if (h.x > 0.0f) {
                    hit = {h.x, i, h.y, h.z};
                    if (any_hit) return hit;
                }
            }
        } else {
            uint32_t left = n.first, right = n.first + 1u;
            if (cnt) cnt->nodes += 2;
            float t_left = slabs(S.nodes[left], o, inv, hit.t);
            float t_right = slabs(S.nodes[right], o, inv, hit.t);
            uint32_t near = t_left <= t_right ? left : right;
            uint32_t far = t_left <= t_right ? right : left;
            if (fmin_(t_left, t_right) < INF) {
                if (fmax_(t_left, t_right) < INF && sp < 64u) stack[sp++] = far;
                current = near;
                descend = true;
            }
        }
        if (!descend) {
            if (sp == 0u) break;
            current = stack[--sp];
        }
    }
    return hit;
}

inline bool occluded(const Scene& S, V3 o, V3 d, float dist, Counters* cnt) {
    return trace(S, o, d, dist, true, cnt).prim != NONE;
}

// ---------------------------------------------------------------- textures
// Bilinear, repeating, mirroring Texture.sample (src/Texture.mjs:164-187)
// and the GPU's filtering sampler; srgb decodes each byte before blending.
float srgb_byte(uint8_t b) {
    float c = (float)b / 255.0f;
    return c <= 0.04045f ? c / 12.92f : std::pow((c + 0.055f) / 1.055f, 2.4f);
}
V3 tex_sample(const uint8_t* data, uint32_t w, uint32_t h, uint32_t layer, V2 uv, bool srgb) {
    float x = uv.x * (float)w - 0.5f, y = uv.y * (float)h - 0.5f;
    float fx0 = std::floor(x), fy0 = std::floor(y);
    float fx = x - fx0, fy = y - fy0;
    int64_t x0 = (int64_t)fx0, y0 = (int64_t)fy0;
    V3 acc = {0.0f, 0.0f, 0.0f};
    const int dxs[4] = {0, 1, 0, 1}, dys[4] = {0, 0, 1, 1};
    const float ws[4] = {(1.0f - fx) * (1.0f - fy), fx * (1.0f - fy), (1.0f - fx) * fy, fx * fy};
    for (int k = 0; k < 4; ++k) {
        int64_t px = ((x0 + dxs[k]) % (int64_t)w + w) % w;
        int64_t py = ((y0 + dys[k]) % (int64_t)h + h) % h;
        const uint8_t* t = data + ((size_t)layer * w * h + (size_t)py * w + px) * 4;
        V3 c = srgb ? v3(srgb_byte(t[0]), srgb_byte(t[1]), srgb_byte(t[2]))
                    : v3(t[0] / 255.0f, t[1] / 255.0f, t[2] / 255.0f);
        acc = acc + c * ws[k];
    }
    return acc;
}

// ---------------------------------------------------------------- surface
// src/shaders.mjs:449-605
struct Surface { V3 p, n, ng; bool front; uint32_t material, light, kind; V2 uv; };

V2 sphere_uv(V3 n, const Prim& p) {
    V3 z = cross(xyz(p.p1), xyz(p.p2));
    float u = std::atan2(-dot(n, z), dot(n, xyz(p.p1))) / TWO_PI;
    return {u - std::floor(u), std::acos(clampf(dot(n, xyz(p.p2)), -1.0f, 1.0f)) / PI};
}

V3 bump(V3 n, V3 ng, V3 tangent, V3 up, V3 t, float strength) {
    V3 right = tangent - n * dot(n, tangent);
    if (length(right) < 1e-9f) return n;
    right = normalize(right);
    V3 above = cross(n, right);
    if (dot(above, up) < 0.0f) above = -above;
    V3 tilted = normalize(right * ((t.x * 2.0f - 1.0f) * strength) + above * ((t.y * 2.0f - 1.0f) * strength) +
                          n * (t.z * 2.0f - 1.0f));
    return dot(tilted, ng) > 0.0f ? tilted : n;
}

Surface surface(const Scene& S, V3 o, V3 d, const Hit& hit) {
    const Prim& p = S.prims[hit.prim];
    V3 point = o + d * hit.t;
    V3 ng, ns, tangent, up;
    V2 uv;
    switch (p.kind) {
    case SPHERE: {
        ng = normalize(point - xyz(p.p0));
        ns = ng;
        uv = sphere_uv(ng, p);
        tangent = cross(xyz(p.p2), ng);
        if (length(tangent) < 1e-6f) tangent = xyz(p.p1);
        up = cross(ng, tangent);
        break;
    }
    case QUAD: {
        ng = normalize(cross(xyz(p.p1), xyz(p.p2)));
        ns = ng;
        uv = {hit.u, 1.0f - hit.v};
        tangent = xyz(p.p1);
        up = xyz(p.p2);
        break;
    }
    case BOX: {
        V3 lp = {dot(xyz(p.p0), point) + p.p0[3], dot(xyz(p.p1), point) + p.p1[3], dot(xyz(p.p2), point) + p.p2[3]};
        V3 a = {std::fabs(lp.x), std::fabs(lp.y), std::fabs(lp.z)};
        V3 half = {1.0f / length(xyz(p.p0)), 1.0f / length(xyz(p.p1)), 1.0f / length(xyz(p.p2))};
        if (a.x >= a.y && a.x >= a.z) {
            ng = xyz(p.p0) * signf(lp.x);
            uv = {(1.0f - signf(lp.x) * lp.z) * half.z, (1.0f - lp.y) * half.y};
            tangent = (-signf(lp.x)) * xyz(p.p2);
            up = xyz(p.p1);
        } else if (a.y >= a.z) {
            ng = xyz(p.p1) * signf(lp.y);
            uv = {(1.0f + lp.x) * half.x, (1.0f + signf(lp.y) * lp.z) * half.z};
            tangent = xyz(p.p0);
            up = (-signf(lp.y)) * xyz(p.p2);
        } else {
            ng = xyz(p.p2) * signf(lp.z);
            uv = {(1.0f + signf(lp.z) * lp.x) * half.x, (1.0f - lp.y) * half.y};
            tangent = signf(lp.z) * xyz(p.p0);
            up = xyz(p.p1);
        }
        ng = normalize(ng);
        ns = ng;
        break;
    }
    default: {
        V3 e1 = xyz(p.p1) - xyz(p.p0), e2 = xyz(p.p2) - xyz(p.p0);
        ng = normalize(cross(e1, e2));
        ns = ng;
        uv = {hit.u, hit.v};
        V2 du1 = {1.0f, 0.0f}, du2 = {0.0f, 1.0f};
        if (p.attributes != NONE) {
            float w = 1.0f - hit.u - hit.v;
            const float* a0 = S.attr + (size_t)p.attributes * 4;
            const float* a1 = a0 + 4;
            const float* a2 = a0 + 8;
            const float* vs = a0 + 12;
            uint32_t flags = to_u32(vs[3]);
            if (flags & 1u) {
                ns = normalize(xyz(a0) * w + xyz(a1) * hit.u + xyz(a2) * hit.v);
                if (dot(ns, ng) < 0.0f) ns = -ns;
            }
            if (flags & 2u) {
                uv = {a0[3] * w + a1[3] * hit.u + a2[3] * hit.v, vs[0] * w + vs[1] * hit.u + vs[2] * hit.v};
                V2 t1 = {a1[3] - a0[3], vs[1] - vs[0]};
                V2 t2 = {a2[3] - a0[3], vs[2] - vs[0]};
                if (std::fabs(t1.x * t2.y - t2.x * t1.y) > 1e-12f) { du1 = t1; du2 = t2; }
            }
        }
        float det = du1.x * du2.y - du2.x * du1.y;
        tangent = (e1 * du2.y - e2 * du1.y) / det;
        up = -((e2 * du1.x - e1 * du2.x) / det);
        break;
    }
    }
    const Mat& m = S.mats[p.material];
    if (m.normal_map != NONE && S.raw->nmap_layers > 0) {
        V2 tuv = {uv.x * m.texture_scale[0], uv.y * m.texture_scale[1]};
        V3 t = tex_sample(S.raw->normal_maps, S.raw->nmap_width, S.raw->nmap_height, m.normal_map, tuv, false);
        ns = bump(ns, ng, tangent, up, t, m.normal_scale);
    }
    bool front = dot(ng, d) < 0.0f;
    float flip = front ? 1.0f : -1.0f;
    return {point, ns * flip, ng * flip, front, p.material, p.light, p.kind, uv};
}

V3 offset(const Surface& s, V3 wi) {
    float eps = 1e-4f * fmax_(1.0f, fmax_(std::fabs(s.p.x), fmax_(std::fabs(s.p.y), std::fabs(s.p.z))));
    return s.p + s.ng * (dot(wi, s.ng) > 0.0f ? eps : -eps);
}

// ---------------------------------------------------------------- sky
// src/shaders.mjs:607-687
struct Ctx {
    Scene S;
    const Params* P;
    const float* emitters = nullptr;
};

V2 env_uv(const Params& P, V3 d) {
    float u = 0.5f + std::atan2(d.x, -d.z) / TWO_PI - P.env_rotation / TWO_PI;
    return {u - std::floor(u), std::acos(clampf(d.y, -1.0f, 1.0f)) / PI};
}

// Bilinear filtering of the half-float map: wraps across u, clamps in v.
V3 env_radiance(const Ctx& C, V3 d) {
    const OracleScene* r = C.S.raw;
    V2 uv = env_uv(*C.P, d);
    uint32_t w = r->env_width, h = r->env_height;
    float x = uv.x * (float)w - 0.5f, y = uv.y * (float)h - 0.5f;
    float fx0 = std::floor(x), fy0 = std::floor(y);
    float fx = x - fx0, fy = y - fy0;
    int64_t x0 = (int64_t)fx0, y0 = (int64_t)fy0;
    V3 acc = {0.0f, 0.0f, 0.0f};
    const int dxs[4] = {0, 1, 0, 1}, dys[4] = {0, 0, 1, 1};
    const float ws[4] = {(1.0f - fx) * (1.0f - fy), fx * (1.0f - fy), (1.0f - fx) * fy, fx * fy};
    for (int k = 0; k < 4; ++k) {
        int64_t px = ((x0 + dxs[k]) % (int64_t)w + w) % w;
        int64_t py = std::min<int64_t>(std::max<int64_t>(y0 + dys[k], 0), (int64_t)h - 1);
        const uint16_t* t = r->env_rgba + ((size_t)py * w + px) * 4;
        acc = acc + v3(half_to_float(t[0]), half_to_float(t[1]), half_to_float(t[2])) * ws[k];
    }
    return acc * C.P->env_strength;
}

inline float env_cdf(const Ctx& C, uint32_t x, uint32_t y) {
    return C.S.raw->env_table[(size_t)y * C.S.raw->table_width + x];
}

uint32_t env_search(const Ctx& C, uint32_t y, uint32_t n, float xi) {
    uint32_t lo = 0u, hi = n - 1u;
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2u;
        if (env_cdf(C, mid, y) > xi) hi = mid;
        else lo = mid + 1u;
    }
    return lo;
}

float env_pdf(const Ctx& C, V3 d) {
    uint32_t cw = C.S.raw->table_width, ch = C.S.raw->table_height;
    V2 uv = env_uv(*C.P, d);
    uint32_t x = std::min(to_u32(uv.x * (float)cw), cw - 1u);
    uint32_t y = std::min(to_u32(uv.y * (float)ch), ch - 1u);
    float row = env_cdf(C, y, ch) - (y > 0u ? env_cdf(C, y - 1u, ch) : 0.0f);
    float cell = env_cdf(C, x, y) - (x > 0u ? env_cdf(C, x - 1u, y) : 0.0f);
    float s = std::sin(uv.y * PI);
    return s > 0.0f ? row * cell * (float)(cw * ch) / (TWO_PI_PI * s) : 0.0f;
}

V3 sample_env(const Ctx& C, Rng& g) {
    uint32_t cw = C.S.raw->table_width, ch = C.S.raw->table_height;
    V2 cell = g.rand2();
    uint32_t y = env_search(C, ch, ch, cell.y);
    uint32_t x = env_search(C, y, cw, cell.x);
    V2 within = g.rand2();
    float u = ((float)x + within.x) / (float)cw + C.P->env_rotation / TWO_PI;
    float v = ((float)y + within.y) / (float)ch;
    float phi = (u - 0.5f) * 2.0f * PI;
    float theta = v * PI;
    return {std::sin(theta) * std::sin(phi), std::cos(theta), -std::sin(theta) * std::cos(phi)};
}

inline bool has_env(const Ctx& C) { return C.P->env_strength > 0.0f; }

V3 sky(const Ctx& C, V3 d) {
    const Params& P = *C.P;
    if (has_env(C)) return env_radiance(C, d);
    if (d.y >= 0.0f) return mix(xyz(P.sky_horizon), xyz(P.sky_zenith), std::sqrt(fmin_(d.y, 1.0f)));
    return mix(xyz(P.sky_horizon), xyz(P.sky_ground), std::sqrt(fmin_(-d.y, 1.0f)));
}

inline bool in_sun(const Params& P, V3 d) {
    return P.sun_on != 0u && dot(d, xyz(P.sun_dir)) >= 1.0f - P.sun_omc;
}

// ---------------------------------------------------------------- materials
// src/shaders.mjs:689-871
V3 albedo(const Ctx& C, const Mat& m, V3 p, V2 uv) {
    V3 c = xyz(m.color);
    if (m.checker_size > 0.0f) {
        V3 q = p / m.checker_size + splat3(1e-3f);
        int32_t cx = to_i32(std::floor(q.x)), cy = to_i32(std::floor(q.y)), cz = to_i32(std::floor(q.z));
        int32_t odd = (int32_t)((uint32_t)cx + (uint32_t)cy + (uint32_t)cz) & 1;
        c = odd != 0 ? xyz(m.checker) : xyz(m.color);
    }
    const OracleScene* r = C.S.raw;
    if (m.texture != NONE && r->tex_layers > 0)
        c = c * tex_sample(r->textures, r->tex_width, r->tex_height, m.texture,
                           {uv.x * m.texture_scale[0], uv.y * m.texture_scale[1]}, true);
    return c;
}

V3 emitted(const Ctx& C, const Mat& m, V2 uv) {
    const OracleScene* r = C.S.raw;
    if (m.emission_texture == NONE || r->tex_layers == 0) return xyz(m.emission);
    return xyz(m.emission) * tex_sample(r->textures, r->tex_width, r->tex_height, m.emission_texture,
                                        {uv.x * m.texture_scale[0], uv.y * m.texture_scale[1]}, true);
}

inline bool is_delta(const Mat& m) { return m.kind == GLASS || (m.kind == METAL && m.roughness < 0.01f); }
inline float ggx_alpha2(float r) { float a = fmax_(r * r, 0.002f); return a * a; }
inline float ggx_d(float nh, float a2) { float k = nh * nh * (a2 - 1.0f) + 1.0f; return a2 / (PI * k * k); }
inline float smith_g1(float nv, float a2) { return 2.0f * nv / (nv + std::sqrt(a2 + (1.0f - a2) * nv * nv)); }

// sample_ggx with its two uniforms given (src/shaders.mjs:728-736).
V3 ggx_half(V3 n, float a2, V2 u) {
    float r = u.x;
    float cos_t = std::sqrt((1.0f - r) / (1.0f + (a2 - 1.0f) * r));
    float sin_t = std::sqrt(fmax_(0.0f, 1.0f - cos_t * cos_t));
    float phi = TWO_PI * u.y;
    return normalize(fmul(frame(n), {sin_t * std::cos(phi), sin_t * std::sin(phi), cos_t}));
}
V3 sample_ggx(Rng& g, V3 n, float a2) { return ggx_half(n, a2, g.rand2()); }

inline V3 schlick(V3 f0, float c) {
    float k = std::pow(1.0f - clampf(c, 0.0f, 1.0f), 5.0f);
    return f0 + (splat3(1.0f) - f0) * k;
}
inline float dielectric_f0(float ior) { float r = (ior - 1.0f) / (ior + 1.0f); return r * r; }

float fresnel_dielectric(float cos_i, float eta) {
    float sin2_t = (1.0f - cos_i * cos_i) / (eta * eta);
    if (sin2_t >= 1.0f) return 1.0f;
    float cos_t = std::sqrt(1.0f - sin2_t);
    float rs = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
    float rp = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
    return 0.5f * (rs * rs + rp * rp);
}

inline float plastic_spec_chance(const Mat& m, float nv) {
    return clampf(schlick(splat3(dielectric_f0(m.ior)), nv).x * 4.0f + 0.1f, 0.1f, 0.9f);
}

struct Eval { V3 f; float pdf; };

Eval eval_bsdf(const Mat& m, V3 base, V3 n, V3 wo, V3 wi) {
    float nl = dot(n, wi), nv = dot(n, wo);
    if (nl <= 0.0f || nv <= 0.0f) return {{0.0f, 0.0f, 0.0f}, 0.0f};
    if (m.kind == DIFFUSE) return {base * (nl / PI), nl / PI};
    float a2 = ggx_alpha2(m.roughness);
    V3 h = normalize(wo + wi);
    float nh = fmax_(dot(n, h), 0.0f);
    float vh = fmax_(dot(wo, h), 1e-6f);
    float d = ggx_d(nh, a2);
    float g = smith_g1(nv, a2) * smith_g1(nl, a2);
    float spec_pdf = d * nh / (4.0f * vh);
    if (m.kind == METAL) {
        V3 f = schlick(base, vh);
        return {f * (d * g / (4.0f * nv)), spec_pdf};
    }
    V3 f0 = splat3(dielectric_f0(m.ior));
    float coat = schlick(f0, vh).x * d * g / (4.0f * nv);
    float through = (1.0f - schlick(f0, nl).x) * (1.0f - schlick(f0, nv).x);
    V3 diffuse = base * (through * nl / PI);
    float chance = plastic_spec_chance(m, nv);
    return {splat3(coat) + diffuse, chance * spec_pdf + (1.0f - chance) * nl / PI};
}

struct BsdfSample { V3 wi, weight; float pdf; bool delta, ok; };

BsdfSample sample_bsdf(Rng& g, const Mat& m, V3 base, const Surface& s, V3 wo) {
    const BsdfSample fail = {{0, 0, 0}, {0, 0, 0}, 0.0f, false, false};
    V3 n = s.n;
    float nv = dot(n, wo);
    if (m.kind == GLASS) {
        float eta = s.front ? m.ior : 1.0f / m.ior;
        float cos_i = clampf(nv, 0.0f, 1.0f);
        float f = fresnel_dielectric(cos_i, eta);
        if (g.rand() < f) return {reflect(-wo, n), {1, 1, 1}, 0.0f, true, true};
        V3 wi = refract(-wo, n, 1.0f / eta);
        if (dot(wi, wi) == 0.0f) return {reflect(-wo, n), {1, 1, 1}, 0.0f, true, true};
        V3 tint = s.front ? base : v3(1.0f, 1.0f, 1.0f);
        return {normalize(wi), tint, 0.0f, true, true};
    }
    if (m.kind == METAL && m.roughness < 0.01f) return {reflect(-wo, n), schlick(base, nv), 0.0f, true, true};
    V3 wi;
    if (m.kind == DIFFUSE) {
        wi = cosine_hemisphere(g, n);
    } else if (m.kind == METAL || g.rand() < plastic_spec_chance(m, nv)) {
        V3 h = sample_ggx(g, n, ggx_alpha2(m.roughness));
        wi = reflect(-wo, h);
    } else {
        wi = cosine_hemisphere(g, n);
    }
    if (dot(wi, s.ng) <= 0.0f) return fail;
    Eval e = eval_bsdf(m, base, n, wo, wi);
    if (e.pdf <= 0.0f) return fail;
    return {wi, e.f / e.pdf, e.pdf, false, true};
}

// ---------------------------------------------------------------- lights
// src/shaders.mjs:873-1018
inline uint32_t light_choices(const Ctx& C) {
    const Params& P = *C.P;
    return P.light_count + (P.sun_on != 0u ? 1u : 0u) + (has_env(C) ? 1u : 0u);
}
inline uint32_t sun_choice(const Params& P) { return P.sun_on != 0u ? P.light_count : NONE; }
inline uint32_t env_choice(const Ctx& C) {
    return has_env(C) ? C.P->light_count + (C.P->sun_on != 0u ? 1u : 0u) : NONE;
}
inline float light_cdf(const Ctx& C, uint32_t l) {
    return bits_f(C.S.lights[C.P->light_count + C.P->volume_count + l]);
}
inline float light_chance(const Ctx& C, uint32_t l) {
    float below = l > 0u ? light_cdf(C, l - 1u) : 0.0f;
    return (light_cdf(C, l) - below) * (float)C.P->light_count / (float)light_choices(C);
}
uint32_t pick_light(const Ctx& C, float u) {
    uint32_t lo = 0u, hi = C.P->light_count - 1u;
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2u;
        if (light_cdf(C, mid) > u) hi = mid;
        else lo = mid + 1u;
    }
    return lo;
}
inline float sphere_omc(float r, float d2) {
    float x = fmin_(r * r / d2, 1.0f);
    return x / (1.0f + std::sqrt(1.0f - x));
}

struct LightSample { V3 wi; float dist; V3 radiance; float pdf; };

LightSample sample_light(const Ctx& C, Rng& g, V3 p, Counters*) {
    const Params& P = *C.P;
    const LightSample none = {{0, 0, 0}, 0.0f, {0, 0, 0}, 0.0f};
    uint32_t choices = light_choices(C);
    float x = g.rand() * (float)choices;
    uint32_t pick = std::min(to_u32(x), choices - 1u);
    float chance = 1.0f / (float)choices;
    if (pick == env_choice(C)) {
        V3 wi = sample_env(C, g);
        float pdf = env_pdf(C, wi);
        if (pdf <= 0.0f) return none;
        return {wi, INF, env_radiance(C, wi), chance * pdf};
    }
    if (pick == sun_choice(P)) {
        V3 wi = sample_cone(g, xyz(P.sun_dir), P.sun_omc);
        return {wi, INF, xyz(P.sun_radiance), chance / (TWO_PI * P.sun_omc)};
    }
    pick = pick_light(C, x / (float)P.light_count);
    chance = light_chance(C, pick);
    const Prim& prim = C.S.prims[C.S.lights[pick]];
    const Mat& m = C.S.mats[prim.material];
    if (prim.kind == QUAD) {
        V2 ab = g.rand2();
        float a = ab.x, b = ab.y;
        V3 q = xyz(prim.p0) + xyz(prim.p1) * a + xyz(prim.p2) * b;
        V3 to = q - p;
        float d2 = dot(to, to);
        float dist = std::sqrt(d2);
        V3 wi = to / dist;
        V3 c = cross(xyz(prim.p1), xyz(prim.p2));
        float area = length(c);
        float cos_l = -dot(wi, c / area);
        if (cos_l <= 0.0f) return none;
        return {wi, dist, emitted(C, m, {a, 1.0f - b}), chance * d2 / (cos_l * area)};
    }
    V3 to = xyz(prim.p0) - p;
    float d2 = dot(to, to);
    float r = prim.p0[3];
    if (d2 <= r * r) return none;
    float omc = sphere_omc(r, d2);
    V3 wi = sample_cone(g, to / std::sqrt(d2), omc);
    float b = dot(to, wi);
    float dist = b - std::sqrt(fmax_(b * b - (d2 - r * r), 0.0f));
    V3 n = normalize(p + wi * dist - xyz(prim.p0));
    return {wi, dist, emitted(C, m, sphere_uv(n, prim)), chance / (TWO_PI * omc)};
}

float light_pdf(const Ctx& C, uint32_t prim_index, V3 p, V3 wi, float t) {
    const Prim& prim = C.S.prims[prim_index];
    float chance = light_chance(C, prim.light);
    if (prim.kind == QUAD) {
        V3 c = cross(xyz(prim.p1), xyz(prim.p2));
        float area = length(c);
        float cos_l = std::fabs(dot(wi, c / area));
        return chance * t * t / fmax_(cos_l * area, 1e-12f);
    }
    V3 to = xyz(prim.p0) - p;
    float omc = sphere_omc(prim.p0[3], dot(to, to));
    return chance / (TWO_PI * omc);
}

inline float power_heuristic(float a, float b) {
    float a2 = a * a, b2 = b * b;
    if (a2 + b2 <= 0.0f) return 0.0f;
    return a2 / (a2 + b2);
}

// ---------------------------------------------------------------- volumes
// src/shaders.mjs:1020-1123
V2 volume_span(const Ctx& C, uint32_t i, V3 o, V3 d) {
    const Prim& p = C.S.prims[i];
    if (p.kind == SPHERE) {
        V3 oc = o - xyz(p.p0);
        float a = dot(d, d);
        float half_b = dot(oc, d);
        float disc = half_b * half_b - a * (dot(oc, oc) - p.p0[3] * p.p0[3]);
        if (disc < 0.0f) return {1.0f, 0.0f};
        float sq = std::sqrt(disc);
        return {fmax_((-half_b - sq) / a, 0.0f), (-half_b + sq) / a};
    }
    V3 lo = {dot(xyz(p.p0), o) + p.p0[3], dot(xyz(p.p1), o) + p.p1[3], dot(xyz(p.p2), o) + p.p2[3]};
    V3 sd = safe({dot(xyz(p.p0), d), dot(xyz(p.p1), d), dot(xyz(p.p2), d)});
    V3 inv = {1.0f / sd.x, 1.0f / sd.y, 1.0f / sd.z};
    V3 t0 = (splat3(-1.0f) - lo) * inv, t1 = (splat3(1.0f) - lo) * inv;
    V3 tmin = vmin(t0, t1), tmax = vmax(t0, t1);
    return {fmax_(fmax_(fmax_(tmin.x, tmin.y), tmin.z), 0.0f), fmin_(fmin_(tmax.x, tmax.y), tmax.z)};
}

inline uint32_t volume(const Ctx& C, uint32_t k) { return C.S.lights[C.P->light_count + k]; }

struct Scatter { float t; uint32_t volume; };

Scatter scatter_in_volumes(const Ctx& C, Rng& g, V3 o, V3 d, float t_max) {
    Scatter s = {t_max, NONE};
    for (uint32_t k = 0; k < C.P->volume_count; k++) {
        uint32_t i = volume(C, k);
        V2 span = volume_span(C, i, o, d);
        float end = fmin_(span.y, s.t);
        if (span.x < end) {
            float t = span.x - std::log(1.0f - g.rand()) / C.S.mats[C.S.prims[i].material].density;
            if (t < end) s = {t, i};
        }
    }
    return s;
}

float transmittance(const Ctx& C, V3 o, V3 d, float dist) {
    float depth = 0.0f;
    for (uint32_t k = 0; k < C.P->volume_count; k++) {
        uint32_t i = volume(C, k);
        V2 span = volume_span(C, i, o, d);
        float end = fmin_(span.y, dist);
        if (span.x < end) depth += C.S.mats[C.S.prims[i].material].density * (end - span.x);
    }
    return std::exp(-depth);
}

inline float henyey_greenstein(float c, float g) {
    float k = 1.0f + g * g - 2.0f * g * c;
    return (1.0f - g * g) / (FOUR_PI * k * std::sqrt(k));
}

V3 sample_henyey_greenstein(Rng& rg, V3 d, float g) {
    V2 u = rg.rand2();
    float c = 1.0f - 2.0f * u.x;
    if (std::fabs(g) >= 1e-3f) {
        float sq = (1.0f - g * g) / (1.0f + g - 2.0f * g * u.x);
        c = clampf((1.0f + g * g - sq * sq) / (2.0f * g), -1.0f, 1.0f);
    }
    float sin_t = std::sqrt(fmax_(0.0f, 1.0f - c * c));
    float phi = TWO_PI * u.y;
    return normalize(fmul(frame(d), {sin_t * std::cos(phi), sin_t * std::sin(phi), c}));
}

// ---------------------------------------------------------------- dispersion
// src/shaders.mjs:1125-1144, 1409-1423
inline float lobe(float x, float mu, float below, float above) {
    float t = (x - mu) / (x < mu ? below : above);
    return std::exp(-0.5f * t * t);
}
V3 wavelength_rgb(float l) {
    float x = 1.056f * lobe(l, 599.8f, 37.9f, 31.0f) + 0.362f * lobe(l, 442.0f, 16.0f, 26.7f) -
              0.065f * lobe(l, 501.1f, 20.4f, 26.2f);
    float y = 0.821f * lobe(l, 568.8f, 46.9f, 40.5f) + 0.286f * lobe(l, 530.9f, 16.3f, 31.1f);
    float z = 1.217f * lobe(l, 437.0f, 11.8f, 36.0f) + 0.681f * lobe(l, 459.0f, 26.0f, 13.8f);
    V3 rgb = {3.2406f * x - 1.5372f * y - 0.4986f * z, -0.9689f * x + 1.8758f * y + 0.0415f * z,
              0.0557f * x - 0.2040f * y + 1.0570f * z};
    return vmax(rgb, splat3(0.0f)) * v3(1.92985f, 2.94689f, 3.11081f);
}
const float INV_D2 = (float)(1.0 / (0.5876 * 0.5876));

Mat dispersed(const Mat& m, float& wavelength, V3& throughput, float u) {
    Mat sm = m;
    if (m.kind == GLASS && m.dispersion > 0.0f) {
        if (wavelength == 0.0f) {
            wavelength = 380.0f + 340.0f * u;
            throughput = throughput * wavelength_rgb(wavelength);
        }
        float l = wavelength * 1e-3f;
        sm.ior = m.ior + m.dispersion * (1.0f / (l * l) - INV_D2);
    }
    return sm;
}

// ---------------------------------------------------------------- paths
// src/shaders.mjs:1148-1407
struct Ray { V3 o, d; };

Ray camera_ray(const Params& P, Rng* g, V2 px, bool jitter) {
    V2 size = {(float)P.width, (float)P.height};
    V2 ndc = {2.0f * px.x / size.x - 1.0f, 1.0f - 2.0f * px.y / size.y};
    V3 dir = normalize(xyz(P.cam_forward) + xyz(P.cam_right) * (ndc.x * P.tan_half_fov * P.aspect) +
                       xyz(P.cam_up) * (ndc.y * P.tan_half_fov));
    if (!jitter || P.lens_radius <= 0.0f) return {xyz(P.cam_pos), dir};
    V3 focus = xyz(P.cam_pos) + dir * (P.focus_distance / dot(dir, xyz(P.cam_forward)));
    V2 lens = concentric_disk(*g);
    lens = {lens.x * P.lens_radius, lens.y * P.lens_radius};
    V3 origin = xyz(P.cam_pos) + xyz(P.cam_right) * lens.x + xyz(P.cam_up) * lens.y;
    return {origin, normalize(focus - origin)};
}

struct PathState {
    V3 g_albedo;
    V4 g_normal;
    float spectrum_u;
};

inline bool roulette(Rng& g, V3& throughput, uint32_t depth) {
    if (depth < 3u) return true;
    V3 t = throughput;
    float survive = fmin_(fmax_(t.x, fmax_(t.y, t.z)), 0.95f);
    if (g.rand() >= survive) return false;
    throughput = t / survive;
    return true;
}

struct PathCount { uint64_t closest = 0, shadow = 0, surfaces = 0, smooth = 0; Counters c; };

V3 radiance(const Ctx& C, Rng& g, PathState& st, V2 px, PathCount* pc) {
    const Params& P = *C.P;
    const Scene& S = C.S;
    Counters* cnt = pc ? &pc->c : nullptr;
    st.g_albedo = {1.0f, 1.0f, 1.0f};
    st.g_normal = {0.0f, 0.0f, 0.0f, 0.0f};
    bool guided = false;
    float travelled = 0.0f;
    Ray ray = camera_ray(P, &g, px, true);
    V3 o = ray.o, d = ray.d;
    V3 throughput = {1.0f, 1.0f, 1.0f};
    V3 result = {0.0f, 0.0f, 0.0f};
    float last_pdf = 0.0f;
    bool last_delta = true;
    V3 last_p = o;
    uint32_t choices = light_choices(C);
    float wavelength = 0.0f;
    bool caustic = false, diffuse_seen = false;

    for (uint32_t depth = 0u; depth <= P.max_bounces; depth++) {
        g.dim = (depth + 1u) << 16u;
        if (pc) pc->closest++;
        Hit hit = trace(S, o, d, INF, false, cnt);

        if (P.volume_count > 0u) {
            Scatter sc = scatter_in_volumes(C, g, o, d, hit.t);
            if (sc.volume != NONE) {
                const Mat& vm = S.mats[S.prims[sc.volume].material];
                V3 p = o + d * sc.t;
                travelled += sc.t;
                if (!guided) {
                    guided = true;
                    st.g_albedo = throughput * xyz(vm.color);
                    V3 nd = -d;
                    st.g_normal = {nd.x, nd.y, nd.z, travelled};
                }
                if (depth == P.max_bounces) break;
                throughput = throughput * xyz(vm.color);
                if (choices > 0u) {
                    LightSample ls = sample_light(C, g, p, cnt);
                    if (ls.pdf > 0.0f) {
                        float phase = henyey_greenstein(dot(d, ls.wi), vm.anisotropy);
                        float reach = ls.dist * (float)(1.0 - 1e-3);
                        if (pc) pc->shadow++;
                        if (!occluded(S, p, ls.wi, reach, cnt))
                            result = result + throughput * ls.radiance *
                                                  (phase * transmittance(C, p, ls.wi, reach) *
                                                   power_heuristic(ls.pdf, phase) / ls.pdf);
                    }
                }
                V3 wi = sample_henyey_greenstein(g, d, vm.anisotropy);
                last_pdf = henyey_greenstein(dot(d, wi), vm.anisotropy);
                last_delta = false;
                last_p = p;
                o = p;
                d = wi;
                caustic = false;
                diffuse_seen = true;
                if (!roulette(g, throughput, depth)) break;
                continue;
            }
        }

        if (hit.prim == NONE) {
            if (!guided) st.g_albedo = throughput;
            float sky_weight = 1.0f;
            if (has_env(C) && !last_delta)
                sky_weight = power_heuristic(last_pdf, env_pdf(C, d) / (float)choices);
            result = result + throughput * sky(C, d) * sky_weight;
            if (in_sun(P, d) && !(caustic && last_delta)) {
                float w = 1.0f;
                if (!last_delta)
                    w = power_heuristic(last_pdf, 1.0f / ((float)choices * 2.0f * PI * P.sun_omc));
                result = result + throughput * xyz(P.sun_radiance) * w;
            }
            break;
        }

        Surface s = surface(S, o, d, hit);
        const Mat& m = S.mats[s.material];
        travelled += hit.t;
        if (pc) {
            pc->surfaces++;
            const Prim& pr = S.prims[hit.prim];
            if (pr.kind == TRIANGLE && pr.attributes != NONE) pc->smooth++;
        }

        if (!guided && (m.kind == LIGHT || !is_delta(m) || depth == P.max_bounces)) {
            guided = true;
            st.g_albedo = throughput * (m.kind == LIGHT ? v3(1.0f, 1.0f, 1.0f) : albedo(C, m, s.p, s.uv));
            st.g_normal = {s.n.x, s.n.y, s.n.z, travelled};
        }

        bool glows = m.emission[0] > 0.0f || m.emission[1] > 0.0f || m.emission[2] > 0.0f;
        if (s.front && glows && !(caustic && last_delta && s.light != NONE)) {
            float w = 1.0f;
            if (!last_delta && s.light != NONE)
                w = power_heuristic(last_pdf, light_pdf(C, hit.prim, last_p, d, hit.t));
            result = result + throughput * emitted(C, m, s.uv) * w;
        }

        if (m.kind == LIGHT) break;
        if (!is_delta(m)) {
            caustic = P.caustics != 0u && !diffuse_seen;
            diffuse_seen = true;
        }
        if (depth == P.max_bounces) break;

        V3 base = albedo(C, m, s.p, s.uv);
        V3 wo = -d;

        if (!is_delta(m) && choices > 0u) {
            LightSample ls = sample_light(C, g, s.p, cnt);
            if (ls.pdf > 0.0f && dot(ls.wi, s.ng) > 0.0f) {
                Eval e = eval_bsdf(m, base, s.n, wo, ls.wi);
                V3 origin = offset(s, ls.wi);
                float reach = (ls.dist - dot(origin - s.p, ls.wi)) * (float)(1.0 - 1e-3);
                if (e.pdf > 0.0f) {
                    if (pc) pc->shadow++;
                    if (!occluded(S, origin, ls.wi, reach, cnt))
                        result = result + throughput * e.f * ls.radiance *
                                              (transmittance(C, origin, ls.wi, reach) *
                                               power_heuristic(ls.pdf, e.pdf) / ls.pdf);
                }
            }
        }

        Mat dm = dispersed(m, wavelength, throughput, st.spectrum_u);
        BsdfSample bs = sample_bsdf(g, dm, base, s, wo);
        if (!bs.ok) break;
        throughput = throughput * bs.weight;
        last_pdf = bs.pdf;
        last_delta = bs.delta;
        last_p = s.p;
        o = offset(s, bs.wi);
        d = bs.wi;
        if (!roulette(g, throughput, depth)) break;
    }
    return result;
}

V3 inspect(const Ctx& C, V2 px) {
    const Params& P = *C.P;
    Ray ray = camera_ray(P, nullptr, px, false);
    Hit hit = trace(C.S, ray.o, ray.d, INF, false);
    if (hit.prim == NONE) {
        if (P.view == VIEW_ALBEDO) return sky(C, ray.d);
        return {0.0f, 0.0f, 0.0f};
    }
    Surface s = surface(C.S, ray.o, ray.d, hit);
    const Mat& m = C.S.mats[s.material];
    if (P.view == VIEW_ALBEDO) {
        if (m.kind == LIGHT) {
            V3 e = emitted(C, m, s.uv);
            return e / fmax_(fmax_(e.x, e.y), fmax_(e.z, 1.0f));
        }
        return albedo(C, m, s.p, s.uv);
    }
    if (P.view == VIEW_NORMAL) return s.n;
    return splat3(hit.t);
}

// One pixel of main(): samples_per_frame paths, accumulated.
void main_pixel(const Ctx& C, uint32_t x, uint32_t y, float* a, float* ga, float* gn, PathCount* pc) {
    const Params& P = *C.P;
    uint32_t index = y * P.width + x;
    V4 sum = {0, 0, 0, 0};
    V3 sum_albedo = {0, 0, 0};
    V4 sum_normal = {0, 0, 0, 0};
    for (uint32_t k = 0; k < P.samples_per_frame; k++) {
        Rng g;
        g.state = pcg(index ^ pcg(P.sample_offset + k));
        g.qmc = true;
        g.seed = pcg(index + 0x9e3779b9u);
        g.index = P.sample_offset + k;
        g.dim = 0u;
        PathState st;
        st.spectrum_u = (float)((pcg(index * 7919u + 1u) + (P.sample_offset + k) * 2654435769u) >> 8u) *
                        (1.0f / 16777216.0f);
        V3 c;
        st.g_albedo = {0, 0, 0};
        st.g_normal = {0, 0, 0, 0};
        if (P.view == VIEW_COLOR) {
            V2 j = g.rand2();
            c = radiance(C, g, st, {(float)x + j.x, (float)y + j.y}, pc);
            float peak = fmax_(c.x, fmax_(c.y, c.z));
            if (P.firefly_clamp > 0.0f && peak > P.firefly_clamp) c = c * (P.firefly_clamp / peak);
        } else {
            c = inspect(C, {(float)x + 0.5f, (float)y + 0.5f});
        }
        if (finite3(c) && finite3(st.g_albedo) && finite3(v3(st.g_normal.x, st.g_normal.y, st.g_normal.z)) &&
            finite3(splat3(st.g_normal.w))) {
            float l = luminance(c);
            sum = {sum.x + c.x, sum.y + c.y, sum.z + c.z, sum.w + l * l};
            sum_albedo = sum_albedo + st.g_albedo;
            sum_normal = {sum_normal.x + st.g_normal.x, sum_normal.y + st.g_normal.y, sum_normal.z + st.g_normal.z,
                          sum_normal.w + st.g_normal.w};
        }
    }
    if (P.sample_offset == 0u) {
        a[0] = sum.x; a[1] = sum.y; a[2] = sum.z; a[3] = sum.w;
        ga[0] = sum_albedo.x; ga[1] = sum_albedo.y; ga[2] = sum_albedo.z; ga[3] = 0.0f;
        gn[0] = sum_normal.x; gn[1] = sum_normal.y; gn[2] = sum_normal.z; gn[3] = sum_normal.w;
    } else {
        a[0] += sum.x; a[1] += sum.y; a[2] += sum.z; a[3] += sum.w;
        ga[0] += sum_albedo.x; ga[1] += sum_albedo.y; ga[2] += sum_albedo.z; ga[3] += 0.0f;
        gn[0] += sum_normal.x; gn[1] += sum_normal.y; gn[2] += sum_normal.z; gn[3] += sum_normal.w;
    }
}

// ---------------------------------------------------------------- light tracing
// src/shaders.mjs:1425-1631, 1630-1831
// Light tracing's outputs, its light paths a frame (`paths`: photon ids
// 0 .. paths - 1, the normalisation of each photon's power) and the photon
// slots there is room for (`capacity`).
struct LightBuffers { uint32_t* splats; uint32_t* heads; float* photons; uint32_t paths, capacity; };

void splat(const Ctx& C, Rng& g, const LightBuffers& B, const Surface& s, const Mat& m, V3 base, V3 d, V3 beta) {
    const Params& P = *C.P;
    V3 to_camera = xyz(P.cam_pos) - s.p;
    float dist = length(to_camera);
    V3 wo = to_camera / dist;
    float cos_film = -dot(wo, xyz(P.cam_forward));
    if (cos_film <= 1e-6f || dot(wo, s.ng) <= 0.0f) return;
    float x = -dot(wo, xyz(P.cam_right)) / (cos_film * P.tan_half_fov * P.aspect);
    float y = -dot(wo, xyz(P.cam_up)) / (cos_film * P.tan_half_fov);
    float px = (x + 1.0f) * 0.5f * (float)P.width;
    float py = (1.0f - y) * 0.5f * (float)P.height;
    if (px < 0.0f || py < 0.0f || px >= (float)P.width || py >= (float)P.height) return;
    V3 wi = -d;
    Eval e = eval_bsdf(m, base, s.n, wo, wi);
    float cos_in = dot(s.n, wi);
    if (e.pdf <= 0.0f || cos_in <= 0.0f) return;
    V3 origin = offset(s, wo);
    if (occluded(C.S, origin, wo, dist - dot(origin - s.p, wo), nullptr)) return;
    float pixel_area = (4.0f * P.tan_half_fov * P.tan_half_fov * P.aspect) / (float)(P.width * P.height);
    V3 value = beta * (e.f / cos_in) * dot(wo, s.ng) / (dist * dist * pixel_area * cos_film * cos_film * cos_film);
    if (!finite3(value)) return;
    uint32_t index = (to_u32(py) * P.width + to_u32(px)) * 4u;
    float r0 = g.rand(), r1 = g.rand(), r2 = g.rand();
    V3 fx = vmin(value * SPLAT_SCALE + v3(r0, r1, r2), splat3(4e9f));
    B.splats[index] += to_u32(fx.x);
    B.splats[index + 1u] += to_u32(fx.y);
    B.splats[index + 2u] += to_u32(fx.z);
}

inline uint32_t photon_count(const Params& P) {
    return std::min(P.width * P.height * P.samples_per_frame, PHOTON_CAPACITY);
}

inline uint32_t photon_cell_hash(int32_t cx, int32_t cy, int32_t cz) {
    return (((uint32_t)cx * 73856093u) ^ ((uint32_t)cy * 19349663u) ^ ((uint32_t)cz * 83492791u)) &
           (PHOTON_TABLE - 1u);
}

void deposit(const Ctx& C, const LightBuffers& B, V3 p, V3 d, V3 beta) {
    uint32_t i = B.heads[PHOTON_TABLE]++;
    if (i >= B.capacity) return;
    V3 q = p / C.P->photon_cell;
    uint32_t h = photon_cell_hash(to_i32(std::floor(q.x)), to_i32(std::floor(q.y)), to_i32(std::floor(q.z)));
    uint32_t next = B.heads[h];
    B.heads[h] = i + 1u;
    float* ph = B.photons + (size_t)i * 12;
    ph[0] = p.x; ph[1] = p.y; ph[2] = p.z; ph[3] = bits_f(next);
    ph[4] = beta.x; ph[5] = beta.y; ph[6] = beta.z; ph[7] = 0.0f;
    ph[8] = -d.x; ph[9] = -d.y; ph[10] = -d.z; ph[11] = 0.0f;
}

V3 gather(const Ctx& C, const uint32_t* heads, const float* photons, const Surface& s, const Mat& m, V3 base, V3 wo) {
    const Params& P = *C.P;
    float r = P.photon_radius, r2 = r * r;
    V3 here = s.p / P.photon_cell;
    int32_t cx = to_i32(std::floor(here.x)), cy = to_i32(std::floor(here.y)), cz = to_i32(std::floor(here.z));
    auto side = [](float h) { return (h - std::floor(h)) >= 0.5f ? 1 : -1; };
    int32_t sx = side(here.x), sy = side(here.y), sz = side(here.z);
    V3 sum = {0, 0, 0};
    uint32_t visited = 0u;
    for (uint32_t k = 0u; k < 8u; k++) {
        int32_t ox = cx + (int32_t)(k & 1u) * sx, oy = cy + (int32_t)((k >> 1u) & 1u) * sy,
                oz = cz + (int32_t)((k >> 2u) & 1u) * sz;
        uint32_t index = heads[photon_cell_hash(ox, oy, oz)];
        for (;;) {
            if (index == 0u || visited >= 4096u) break;
            visited++;
            const float* at = photons + (size_t)(index - 1u) * 12;
            V3 off = xyz(at) - s.p;
            if (dot(off, off) < r2 && std::fabs(dot(off, s.ng)) < 0.25f * r) {
                V3 wi = xyz(at + 8);
                float cos_in = dot(s.n, wi);
                if (cos_in > 0.0f && dot(s.ng, wi) > 0.0f) {
                    Eval e = eval_bsdf(m, base, s.n, wo, wi);
                    sum = sum + xyz(at + 4) * (e.f / cos_in);
                }
            }
            index = f_bits(at[3]);
        }
    }
    return sum / (PI * r2);
}

void light_trace_one(const Ctx& C, const LightBuffers& B, uint32_t id) {
    const Params& P = *C.P;
    const Scene& S = C.S;
    uint32_t photons = B.paths;
    if (id >= photons) return;
    Rng g;
    g.state = pcg(id ^ pcg(P.sample_offset * 2654435769u + 977u));
    const float* E = C.emitters;
    uint32_t sources = f_bits(E[0]), targets = f_bits(E[1]);
    float reach = E[2], circles = E[3];
    float xi = g.rand();
    uint32_t k = 0u;
    for (;;) {
        if (k + 1u >= sources || E[(1u + k) * 4] > xi) break;
        k++;
    }
    const float* entry = E + (1u + k) * 4;
    float chance = entry[2];
    if (chance <= 0.0f) return;
    uint32_t prim_index = f_bits(entry[1]);
    uint32_t first_target = 1u + sources;
    auto tgt = [&](uint32_t i) { return E + (first_target + i) * 4; };
    V3 o, d, beta;
    if (prim_index == NONE) {
        d = -sample_cone(g, xyz(P.sun_dir), P.sun_omc);
        float pick = g.rand() * circles / PI;
        uint32_t j = 0u;
        for (;;) {
            float r = tgt(j)[3];
            if (j + 1u >= targets || pick < r * r) break;
            pick -= r * r;
            j++;
        }
        const float* circle = tgt(j);
        Frame f = frame(d);
        V2 disk = concentric_disk(g);
        disk = {disk.x * circle[3], disk.y * circle[3]};
        o = xyz(circle) - d * reach + f.c0 * disk.x + f.c1 * disk.y;
        float through = 1.0f;
        for (uint32_t i = 0u; i < targets; i++) {
            const float* c = tgt(i);
            V3 q = o - xyz(c);
            V3 across = q - d * dot(q, d);
            through += (i != j && dot(across, across) < c[3] * c[3]) ? 1.0f : 0.0f;
        }
        V3 irradiance = xyz(P.sun_radiance) * (TWO_PI * P.sun_omc);
        beta = irradiance * circles / (through * chance * (float)photons);
    } else {
        const Prim& prim = S.prims[prim_index];
        const Mat& m = S.mats[prim.material];
        V3 n;
        float area;
        V2 uv;
        if (prim.kind == QUAD) {
            float a = g.rand();
            float b = g.rand();
            o = xyz(prim.p0) + xyz(prim.p1) * a + xyz(prim.p2) * b;
            V3 c = cross(xyz(prim.p1), xyz(prim.p2));
            area = length(c);
            n = c / area;
            uv = {a, 1.0f - b};
        } else {
            float z = 1.0f - 2.0f * g.rand();
            float r = std::sqrt(fmax_(0.0f, 1.0f - z * z));
            float phi = TWO_PI * g.rand();
            n = {r * std::cos(phi), z, r * std::sin(phi)};
            o = xyz(prim.p0) + n * prim.p0[3];
            area = FOUR_PI * prim.p0[3] * prim.p0[3];
            uv = sphere_uv(n, prim);
        }
        float cones = 0.0f;
        bool inside = false;
        for (uint32_t i = 0u; i < targets; i++) {
            const float* aim = tgt(i);
            V3 to = xyz(aim) - o;
            float d2 = dot(to, to);
            inside = inside || d2 <= aim[3] * aim[3];
            cones += sphere_omc(aim[3], fmax_(d2, aim[3] * aim[3]));
        }
        if (!inside) {
            float pick = g.rand() * cones;
            uint32_t j = 0u;
            for (;;) {
                const float* aim = tgt(j);
                V3 to = xyz(aim) - o;
                float omc = sphere_omc(aim[3], dot(to, to));
                if (j + 1u >= targets || pick < omc) {
                    d = sample_cone(g, to / length(to), omc);
                    break;
                }
                pick -= omc;
                j++;
            }
            float within = 1.0f;
            for (uint32_t i = 0u; i < targets; i++) {
                const float* aim = tgt(i);
                V3 to = xyz(aim) - o;
                float omc = sphere_omc(aim[3], dot(to, to));
                within += (i != j && 1.0f - dot(d, to / length(to)) <= omc) ? 1.0f : 0.0f;
            }
            float cos_out = dot(d, n);
            if (cos_out <= 0.0f) return;
            beta = emitted(C, m, uv) * (cos_out * area * 2.0f * PI * cones) / (within * chance * (float)photons);
        } else {
            d = cosine_hemisphere(g, n);
            beta = emitted(C, m, uv) * (PI * area) / (chance * (float)photons);
        }
        o = o + n * (1e-4f * fmax_(1.0f, fmax_(std::fabs(o.x), fmax_(std::fabs(o.y), std::fabs(o.z)))));
    }

    float wavelength = 0.0f;
    uint32_t bounces = 0u;
    V3 carried = {1.0f, 1.0f, 1.0f};
    for (uint32_t depth = 0u; depth < P.max_bounces; depth++) {
        Hit hit = trace(S, o, d, INF, false);
        if (hit.prim == NONE) return;
        Surface s = surface(S, o, d, hit);
        const Mat& m = S.mats[s.material];
        if (m.kind == LIGHT) return;
        V3 base = albedo(C, m, s.p, s.uv);
        if (!is_delta(m)) {
            if (bounces > 0u) {
                splat(C, g, B, s, m, base, d, beta * carried);
                deposit(C, B, s.p, d, beta * carried);
            }
            return;
        }
        float u = g.rand();
        Mat dm = dispersed(m, wavelength, carried, u);
        BsdfSample bs = sample_bsdf(g, dm, base, s, -d);
        if (!bs.ok) return;
        carried = carried * bs.weight;
        bounces++;
        o = offset(s, bs.wi);
        d = bs.wi;
        if (!roulette(g, carried, depth)) return;
    }
}

void photon_gather_pixel(const Ctx& C, const uint32_t* heads, const float* photons, uint32_t x, uint32_t y,
                         float* accum) {
    const Params& P = *C.P;
    const Scene& S = C.S;
    uint32_t index = y * P.width + x;
    Rng g;
    g.state = pcg(index ^ pcg(P.sample_offset * 2654435769u + 4099u));
    float jx = g.rand();
    float jy = g.rand();
    Ray ray = camera_ray(P, &g, {(float)x + jx, (float)y + jy}, true);
    V3 o = ray.o, d = ray.d;
    V3 throughput = {1, 1, 1};
    for (uint32_t depth = 0u; depth < P.max_bounces; depth++) {
        Hit hit = trace(S, o, d, INF, false);
        if (hit.prim == NONE) return;
        Surface s = surface(S, o, d, hit);
        const Mat& m = S.mats[s.material];
        if (m.kind == LIGHT) return;
        V3 base = albedo(C, m, s.p, s.uv);
        if (!is_delta(m)) {
            if (depth > 0u) {
                V3 light = throughput * gather(C, heads, photons, s, m, base, -d);
                if (finite3(light)) {
                    float l = luminance(light);
                    float spp = (float)P.samples_per_frame;
                    float* a = accum + (size_t)index * 4;
                    a[0] += light.x * spp; a[1] += light.y * spp; a[2] += light.z * spp; a[3] += l * l * spp;
                }
            }
            return;
        }
        BsdfSample bs = sample_bsdf(g, m, base, s, -d);
        if (!bs.ok) return;
        throughput = throughput * bs.weight;
        o = offset(s, bs.wi);
        d = bs.wi;
    }
}

// ---------------------------------------------------------------- denoiser
// src/shaders.mjs:1916-2098
const float KERNEL[3] = {0.375f, 0.25f, 0.0625f};

struct Den {
    const Params* P;
    const float* accum;
    const float* galb;
    const float* gnorm;
    float samples() const { return (float)std::max(P->total_samples, 1u); }
    V3 albedo_at(uint32_t i) const {
        const float* g = galb + (size_t)i * 4;
        return vmax(v3(g[0], g[1], g[2]) / samples(), splat3(0.01f));
    }
    V4 normal_at(uint32_t i) const {
        const float* g = gnorm + (size_t)i * 4;
        V3 n = xyz(g);
        V3 nn = dot(n, n) > 0.0f ? normalize(n) : v3(0, 0, 0);
        return {nn.x, nn.y, nn.z, g[3] / samples()};
    }
    V3 light_at(uint32_t i) const {
        const float* a = accum + (size_t)i * 4;
        return (v3(a[0], a[1], a[2]) / samples()) / albedo_at(i);
    }
    bool inside(int32_t x, int32_t y) const {
        return x >= 0 && y >= 0 && x < (int32_t)P->width && y < (int32_t)P->height;
    }
};

float similarity(V4 p, V4 q, float step) {
    if (p.w <= 0.0f || q.w <= 0.0f) return (p.w <= 0.0f && q.w <= 0.0f) ? 1.0f : 0.0f;
    if (p.x * p.x + p.y * p.y + p.z * p.z == 0.0f) return 0.0f;
    float wn = std::pow(fmax_(dot(v3(p.x, p.y, p.z), v3(q.x, q.y, q.z)), 0.0f), 64.0f);
    float wz = std::exp(-std::fabs(p.w - q.w) / (p.w * 0.02f * step + 1e-6f));
    return wn * wz;
}

// ---------------------------------------------------------------- FP64 raycast
// The reference's CPU raycast in double precision over the FP32-packed data
// (src/Scene.mjs:202-253, 549-674): reciprocal 1/d with no safe(), NaN slabs
// skipped, popped nodes re-tested against the best t, nearer child first.
struct D3 { double x, y, z; };
inline D3 d3(const float* p) { return {p[0], p[1], p[2]}; }
inline D3 dsub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 dcross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

double slabs64(const Node& n, const double* o, const double* inv) {
    double t_near = 0.0, t_far = INFINITY;
    for (int a = 0; a < 3; ++a) {
        double t0 = ((double)n.mn[a] - o[a]) * inv[a], t1 = ((double)n.mx[a] - o[a]) * inv[a];
        if (std::isnan(t0) || std::isnan(t1)) continue;
        if (t0 > t1) std::swap(t0, t1);
        t_near = std::max(t_near, t0);
        t_far = std::min(t_far, t1);
    }
    return t_near <= t_far ? t_near : INFINITY;
}

bool intersect64(const Prim& p, const double* o3, const double* d3_, double t_max, double* out) {
    const double TM = 1e-4;
    D3 o = {o3[0], o3[1], o3[2]}, d = {d3_[0], d3_[1], d3_[2]};
    if (p.kind == SPHERE) {
        D3 oc = dsub(o, d3(p.p0));
        double r = p.p0[3], a = ddot(d, d), hb = ddot(oc, d), c = ddot(oc, oc) - r * r;
        double disc = hb * hb - a * c;
        if (disc < 0.0) return false;
        double sq = std::sqrt(disc), t = (-hb - sq) / a;
        if (t <= TM || t >= t_max) {
            t = (-hb + sq) / a;
            if (t <= TM || t >= t_max) return false;
        }
        out[0] = t; out[1] = 0.0; out[2] = 0.0;
        return true;
    }
    if (p.kind == QUAD) {
        D3 q = d3(p.p0), qu = d3(p.p1), qv = d3(p.p2), n = dcross(qu, qv);
        double denom = ddot(n, d);
        if (std::fabs(denom) < 1e-12) return false;
        double t = ddot(n, dsub(q, o)) / denom;
        if (t <= TM || t >= t_max) return false;
        D3 h = {o.x + d.x * t - q.x, o.y + d.y * t - q.y, o.z + d.z * t - q.z};
        double nn = ddot(n, n), al = ddot(n, dcross(h, qv)) / nn, be = ddot(n, dcross(qu, h)) / nn;
        if (al < 0.0 || al > 1.0 || be < 0.0 || be > 1.0) return false;
        out[0] = t; out[1] = al; out[2] = be;
        return true;
    }
    if (p.kind == BOX) {
        const float* rows[3] = {p.p0, p.p1, p.p2};
        double t_near = -INFINITY, t_far = INFINITY;
        for (int a = 0; a < 3; ++a) {
            const float* r = rows[a];
            double lo = (double)r[0] * o.x + (double)r[1] * o.y + (double)r[2] * o.z + (double)r[3];
            double ld = (double)r[0] * d.x + (double)r[1] * d.y + (double)r[2] * d.z;
            double inv = 1.0 / ld, t0 = (-1.0 - lo) * inv, t1 = (1.0 - lo) * inv;
            if (std::isnan(t0) || std::isnan(t1)) continue;
            if (t0 > t1) std::swap(t0, t1);
            t_near = std::max(t_near, t0);
            t_far = std::min(t_far, t1);
        }
        if (t_near > t_far) return false;
        double t = t_near > TM ? t_near : t_far;
        if (t <= TM || t >= t_max) return false;
        out[0] = t; out[1] = 0.0; out[2] = 0.0;
        return true;
    }
    if (p.kind != TRIANGLE) return false;
    D3 v0 = d3(p.p0), e1 = dsub(d3(p.p1), v0), e2 = dsub(d3(p.p2), v0), pv = dcross(d, e2);
    double det = ddot(e1, pv);
    if (std::fabs(det) < 1e-12) return false;
    double id = 1.0 / det;
    D3 tv = dsub(o, v0);
    double bu = ddot(tv, pv) * id;
    if (bu < 0.0 || bu > 1.0) return false;
    D3 qv = dcross(tv, e1);
    double bv = ddot(d, qv) * id;
    if (bv < 0.0 || bu + bv > 1.0) return false;
    double t = ddot(e2, qv) * id;
    if (t <= TM || t >= t_max) return false;
    out[0] = t; out[1] = bu; out[2] = bv;
    return true;
}

void trace64(const Scene& S, const double* o, const double* d, double t_max, double* out) {
    double inv[3] = {1.0 / d[0], 1.0 / d[1], 1.0 / d[2]};
    double best_t = t_max, bu = 0, bv = 0;
    int64_t best = -1;
    std::vector<uint32_t> stack = {0};
    while (!stack.empty()) {
        uint32_t n = stack.back();
        stack.pop_back();
        const Node& node = S.nodes[n];
        if (slabs64(node, o, inv) >= best_t) continue;
        if (node.count == 0u) {
            double tl = slabs64(S.nodes[node.first], o, inv), tr = slabs64(S.nodes[node.first + 1], o, inv);
            if (tl <= tr) { stack.push_back(node.first + 1); stack.push_back(node.first); }
            else { stack.push_back(node.first); stack.push_back(node.first + 1); }
            continue;
        }
        for (uint32_t p = node.first; p < node.first + node.count; ++p) {
            double h[3];
            if (intersect64(S.prims[p], o, d, best_t, h)) { best_t = h[0]; bu = h[1]; bv = h[2]; best = p; }
        }
    }
    out[0] = best < 0 ? t_max : best_t;
    out[1] = best < 0 ? -1.0 : (double)best;
    out[2] = bu;
    out[3] = bv;
}

}  // namespace

static void run_threads(uint32_t n, int threads, const std::function<void(uint32_t, uint32_t)>& body) {
    if (threads <= 1 || n < 64) {
        body(0, n);
        return;
    }
    std::vector<std::thread> pool;
    uint32_t chunk = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        uint32_t a = std::min(n, t * chunk), b = std::min(n, a + chunk);
        if (a < b) pool.emplace_back(body, a, b);
    }
    for (auto& th : pool) th.join();
}

// ================================================================= C API
extern "C" {

uint32_t oracle_pcg(uint32_t v) { return pcg(v); }
uint32_t oracle_owen(uint32_t x, uint32_t seed) { return owen(x, seed); }
uint32_t oracle_sobol2(uint32_t i) { return sobol2(i); }
uint32_t oracle_reverse_bits(uint32_t x) { return reverse_bits(x); }

void oracle_qmc_draws(uint32_t pixel, uint32_t sample, uint32_t dim0, uint32_t n, int pairs, float* out) {
    Rng g;
    g.qmc = true;
    g.seed = pcg(pixel + 0x9e3779b9u);
    g.index = sample;
    g.dim = dim0;
    for (uint32_t i = 0; i < n; ++i) {
        if (pairs) {
            V2 r = g.rand2();
            out[2 * i] = r.x;
            out[2 * i + 1] = r.y;
        } else {
            out[i] = g.rand();
        }
    }
}

void oracle_trace(const OracleScene* s, const float* rays, uint32_t n, int any_hit, uint32_t* hits,
                  uint32_t* counters, int threads) {
    Scene S = view(s);
    run_threads(n, threads, [&](uint32_t a, uint32_t b) {
        for (uint32_t i = a; i < b; ++i) {
            const float* r = rays + (size_t)i * 8;
            Counters c;
            Hit h = trace(S, {r[0], r[1], r[2]}, {r[4], r[5], r[6]}, r[3], any_hit != 0, counters ? &c : nullptr);
            hits[(size_t)i * 4] = f_bits(h.t);
            hits[(size_t)i * 4 + 1] = h.prim;
            hits[(size_t)i * 4 + 2] = f_bits(h.u);
            hits[(size_t)i * 4 + 3] = f_bits(h.v);
            if (counters) {
                counters[(size_t)i * 2] = (uint32_t)c.nodes;
                counters[(size_t)i * 2 + 1] = (uint32_t)c.prims;
            }
        }
    });
}

void oracle_trace64(const OracleScene* s, const float* rays, uint32_t n, double* hits, int threads) {
    Scene S = view(s);
    run_threads(n, threads, [&](uint32_t a, uint32_t b) {
        for (uint32_t i = a; i < b; ++i) {
            const float* r = rays + (size_t)i * 8;
            double o[3] = {r[0], r[1], r[2]}, d[3] = {r[4], r[5], r[6]};
            trace64(S, o, d, r[3] >= 1e30f ? INFINITY : (double)r[3], hits + (size_t)i * 4);
        }
    });
}

void oracle_brute_force(const OracleScene* s, const float* rays, uint32_t n, uint32_t prim_count, uint32_t* hits) {
    Scene S = view(s);
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = rays + (size_t)i * 8;
        float best = r[3];
        uint32_t best_p = NONE;
        for (uint32_t k = 0; k < prim_count; ++k) {
            V3 h = intersect(S, k, {r[0], r[1], r[2]}, {r[4], r[5], r[6]}, INF);
            if (h.x > 0.0f && h.x < best) {
                best = h.x;
                best_p = k;
            }
        }
        hits[(size_t)i * 2] = f_bits(best);
        hits[(size_t)i * 2 + 1] = best_p;
    }
}

void oracle_camera_rays(const uint8_t* params, float* rays) {
    const Params& P = *reinterpret_cast<const Params*>(params);
    for (uint32_t y = 0; y < P.height; ++y)
        for (uint32_t x = 0; x < P.width; ++x) {
            Ray r = camera_ray(P, nullptr, {(float)x + 0.5f, (float)y + 0.5f}, false);
            float* o = rays + ((size_t)y * P.width + x) * 8;
            o[0] = r.o.x; o[1] = r.o.y; o[2] = r.o.z; o[3] = INF;
            o[4] = r.d.x; o[5] = r.d.y; o[6] = r.d.z; o[7] = 0.0f;
        }
}

void oracle_main(const OracleScene* s, const uint8_t* params, float* accum, float* guide_albedo,
                 float* guide_normal, uint32_t y0, uint32_t y1, int threads) {
    Ctx C{view(s), reinterpret_cast<const Params*>(params)};
    const uint32_t W = C.P->width;
    run_threads(y1 - y0, threads, [&](uint32_t a, uint32_t b) {
        for (uint32_t y = y0 + a; y < y0 + b; ++y)
            for (uint32_t x = 0; x < W; ++x) {
                size_t o = ((size_t)y * W + x) * 4;
                main_pixel(C, x, y, accum + o, guide_albedo + o, guide_normal + o, nullptr);
            }
    });
}

void oracle_path_stats(const OracleScene* s, const uint8_t* params, uint32_t stride, uint64_t* counts) {
    Ctx C{view(s), reinterpret_cast<const Params*>(params)};
    const Params& P = *C.P;
    PathCount pc;
    uint64_t samples = 0;
    float a[4] = {0, 0, 0, 0}, ga[4] = {0, 0, 0, 0}, gn[4] = {0, 0, 0, 0};
    for (uint32_t i = 0; i < P.width * P.height; i += stride) {
        main_pixel(C, i % P.width, i / P.width, a, ga, gn, &pc);
        samples += P.samples_per_frame;
    }
    counts[0] = pc.closest;
    counts[1] = pc.shadow;
    counts[2] = pc.c.nodes;
    counts[3] = pc.c.prims;
    counts[4] = pc.surfaces;
    counts[5] = pc.smooth;
    counts[6] = samples;
}

void oracle_light_trace(const OracleScene* s, const uint8_t* params, const float* emitters, uint32_t first,
                        uint32_t count, uint32_t* splats, uint32_t* photon_heads, float* photons) {
    const Params& P = *reinterpret_cast<const Params*>(params);
    oracle_light_trace_budget(s, params, emitters, first, count, photon_count(P), PHOTON_CAPACITY, splats,
                              photon_heads, photons);
}

void oracle_light_trace_budget(const OracleScene* s, const uint8_t* params, const float* emitters, uint32_t first,
                               uint32_t count, uint32_t paths, uint32_t capacity, uint32_t* splats,
                               uint32_t* photon_heads, float* photons) {
    Ctx C{view(s), reinterpret_cast<const Params*>(params), emitters};
    LightBuffers B{splats, photon_heads, photons, paths, capacity};
    for (uint32_t id = first; id < first + count; ++id) light_trace_one(C, B, id);
}

void oracle_resolve_splats(const uint8_t* params, uint32_t* splats, float* accum) {
    const Params& P = *reinterpret_cast<const Params*>(params);
    for (uint32_t i = 0; i < P.width * P.height; ++i) {
        V3 fixed = {(float)splats[i * 4], (float)splats[i * 4 + 1], (float)splats[i * 4 + 2]};
        splats[i * 4] = splats[i * 4 + 1] = splats[i * 4 + 2] = 0u;
        V3 light = fixed / SPLAT_SCALE;
        float l = luminance(light);
        float spp = (float)P.samples_per_frame;
        float* a = accum + (size_t)i * 4;
        a[0] += light.x * spp; a[1] += light.y * spp; a[2] += light.z * spp; a[3] += l * l * spp;
    }
}

void oracle_photon_gather(const OracleScene* s, const uint8_t* params, const uint32_t* photon_heads,
                          const float* photons, float* accum, int threads) {
    Ctx C{view(s), reinterpret_cast<const Params*>(params)};
    const uint32_t W = C.P->width;
    run_threads(C.P->height, threads, [&](uint32_t a, uint32_t b) {
        for (uint32_t y = a; y < b; ++y)
            for (uint32_t x = 0; x < W; ++x) photon_gather_pixel(C, photon_heads, photons, x, y, accum);
    });
}

void oracle_prepare(const uint8_t* params, const float* accum, const float* guide_albedo, const float* guide_normal,
                    float* light_out) {
    Den D{reinterpret_cast<const Params*>(params), accum, guide_albedo, guide_normal};
    const Params& P = *D.P;
    for (uint32_t y = 0; y < P.height; ++y)
        for (uint32_t x = 0; x < P.width; ++x) {
            uint32_t i = y * P.width + x;
            float n = D.samples();
            V3 light = D.light_at(i);
            V4 gp = D.normal_at(i);
            float a = luminance(D.albedo_at(i));
            float variance;
            if (n >= 4.0f) {
                const float* ac = accum + (size_t)i * 4;
                float mean = luminance(xyz(ac)) / n;
                float spread = fmax_(ac[3] / n - mean * mean, 0.0f) * n / (n - 1.0f);
                variance = spread / n / (a * a);
            } else {
                float sum = 0.0f, sum2 = 0.0f, weight = 0.0f;
                for (int32_t dy = -2; dy <= 2; dy++)
                    for (int32_t dx = -2; dx <= 2; dx++) {
                        int32_t xx = (int32_t)x + dx, yy = (int32_t)y + dy;
                        if (!D.inside(xx, yy)) continue;
                        uint32_t j = (uint32_t)yy * P.width + (uint32_t)xx;
                        float w = similarity(gp, D.normal_at(j), 1.0f);
                        float l = luminance(D.light_at(j));
                        sum += w * l;
                        sum2 += w * l * l;
                        weight += w;
                    }
                float mean = sum / fmax_(weight, 1e-6f);
                variance = fmax_(sum2 / fmax_(weight, 1e-6f) - mean * mean, 0.0f) + 0.01f * mean * mean;
            }
            float* o = light_out + (size_t)i * 4;
            o[0] = light.x; o[1] = light.y; o[2] = light.z; o[3] = variance;
        }
}

void oracle_atrous(const uint8_t* params, uint32_t step_u, const float* guide_normal, const float* light_in,
                   float* light_out) {
    Den D{reinterpret_cast<const Params*>(params), nullptr, nullptr, guide_normal};
    const Params& P = *D.P;
    int32_t step = (int32_t)step_u;
    for (uint32_t y = 0; y < P.height; ++y)
        for (uint32_t x = 0; x < P.width; ++x) {
            uint32_t i = y * P.width + x;
            const float* centre = light_in + (size_t)i * 4;
            V4 gp = D.normal_at(i);
            float lp = luminance(xyz(centre));
            float variance = 0.0f, variance_weight = 0.0f;
            for (int32_t dy = -1; dy <= 1; dy++)
                for (int32_t dx = -1; dx <= 1; dx++) {
                    int32_t xx = (int32_t)x + dx, yy = (int32_t)y + dy;
                    if (D.inside(xx, yy)) {
                        float w = KERNEL[std::abs(dx)] * KERNEL[std::abs(dy)];
                        variance += w * light_in[((size_t)yy * P.width + xx) * 4 + 3];
                        variance_weight += w;
                    }
                }
            float sigma = 4.0f * std::sqrt(variance / variance_weight) + 1e-4f;
            V3 sum = {0, 0, 0};
            float sum_variance = 0.0f, weight = 0.0f;
            for (int32_t dy = -2; dy <= 2; dy++)
                for (int32_t dx = -2; dx <= 2; dx++) {
                    int32_t xx = (int32_t)x + dx * step, yy = (int32_t)y + dy * step;
                    if (!D.inside(xx, yy)) continue;
                    uint32_t j = (uint32_t)yy * P.width + (uint32_t)xx;
                    const float* q = light_in + (size_t)j * 4;
                    float w = KERNEL[std::abs(dx)] * KERNEL[std::abs(dy)] * similarity(gp, D.normal_at(j), (float)step) *
                              std::exp(-std::fabs(lp - luminance(xyz(q))) / sigma);
                    sum = sum + w * xyz(q);
                    sum_variance += w * w * q[3];
                    weight += w;
                }
            float* o = light_out + (size_t)i * 4;
            if (weight > 0.0f) {
                V3 r = sum / weight;
                o[0] = r.x; o[1] = r.y; o[2] = r.z; o[3] = sum_variance / (weight * weight);
            } else {
                o[0] = centre[0]; o[1] = centre[1]; o[2] = centre[2]; o[3] = centre[3];
            }
        }
}

void oracle_compose(const uint8_t* params, const float* guide_albedo, const float* light_in, float* out) {
    Den D{reinterpret_cast<const Params*>(params), nullptr, guide_albedo, nullptr};
    const Params& P = *D.P;
    for (uint32_t i = 0; i < P.width * P.height; ++i) {
        V3 r = xyz(light_in + (size_t)i * 4) * D.albedo_at(i);
        float* o = out + (size_t)i * 4;
        o[0] = r.x; o[1] = r.y; o[2] = r.z; o[3] = 0.0f;
    }
}

// DISPLAY_SHADER's fragment (src/shaders.mjs:2100-2152): the averaged
// accumulation (or the denoised average), exposed, Narkowicz's ACES fit and
// sRGB for the colour view; sRGB for albedo; n·0.5+0.5 for normals; depth as
// 1/(1 + d/focus). Written as RGBA8, each channel round(clamp(x, 0, 1)·255)
// as an 8-bit unorm target stores it.
static float aces1(float x) {
    float v = (x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f);
    return std::fmin(std::fmax(v, 0.0f), 1.0f);
}
static float srgb1(float c) {
    float x = std::fmin(std::fmax(c, 0.0f), 1.0f);
    return x <= 0.0031308f ? x * 12.92f : 1.055f * std::pow(x, 1.0f / 2.4f) - 0.055f;
}
static uint8_t unorm8(float x) {
    float c = std::fmin(std::fmax(x, 0.0f), 1.0f);
    return (uint8_t)std::nearbyint(c * 255.0f);
}

void oracle_display(const uint8_t* params, const float* accum, const float* denoised, uint8_t* out) {
    const Params& P = *reinterpret_cast<const Params*>(params);
    for (uint32_t i = 0; i < P.width * P.height; ++i) {
        uint8_t* o = out + (size_t)i * 4;
        float rgb[3] = {0.0f, 0.0f, 0.0f};
        if (P.total_samples != 0u) {
            const float n = (float)P.total_samples;
            V3 c = xyz(accum + (size_t)i * 4) / n;
            if (P.view == 0u) {
                V3 shown = P.denoise != 0u ? xyz(denoised + (size_t)i * 4) : c;
                V3 e = shown * P.exposure;
                rgb[0] = srgb1(aces1(e.x)); rgb[1] = srgb1(aces1(e.y)); rgb[2] = srgb1(aces1(e.z));
            } else if (P.view == 1u) {
                rgb[0] = srgb1(c.x); rgb[1] = srgb1(c.y); rgb[2] = srgb1(c.z);
            } else if (P.view == 2u) {
                rgb[0] = c.x * 0.5f + 0.5f; rgb[1] = c.y * 0.5f + 0.5f; rgb[2] = c.z * 0.5f + 0.5f;
            } else {
                float g = c.x <= 0.0f ? 0.0f : 1.0f / (1.0f + c.x / P.focus_distance);
                rgb[0] = rgb[1] = rgb[2] = g;
            }
        }
        o[0] = unorm8(rgb[0]); o[1] = unorm8(rgb[1]); o[2] = unorm8(rgb[2]); o[3] = 255;
    }
}

void oracle_intersect(const OracleScene* s, uint32_t prim, const float* ray, float t_max, float* out) {
    Scene S = view(s);
    V3 h = intersect(S, prim, {ray[0], ray[1], ray[2]}, {ray[4], ray[5], ray[6]}, t_max);
    out[0] = h.x; out[1] = h.y; out[2] = h.z;
}

void oracle_eval_bsdf(const uint32_t* material, const float* base, const float* n, const float* wo, const float* wi,
                      float* out4) {
    const Mat& m = *reinterpret_cast<const Mat*>(material);
    Eval e = eval_bsdf(m, xyz(base), xyz(n), xyz(wo), xyz(wi));
    out4[0] = e.f.x; out4[1] = e.f.y; out4[2] = e.f.z; out4[3] = e.pdf;
}

float oracle_fresnel_dielectric(float cos_i, float eta) { return fresnel_dielectric(cos_i, eta); }

void oracle_ggx_terms(float roughness, float nh, float nv, float nl, float* out3) {
    const float a2 = ggx_alpha2(roughness);
    out3[0] = ggx_d(nh, a2);
    out3[1] = smith_g1(nv, a2);
    out3[2] = smith_g1(nl, a2);
}

void oracle_ggx_sample(float roughness, float u1, float u2, float* h3) {
    V3 h = ggx_half({0.0f, 0.0f, 1.0f}, ggx_alpha2(roughness), {u1, u2});
    h3[0] = h.x; h3[1] = h.y; h3[2] = h.z;
}

void oracle_sample_bsdf(const uint32_t* material, const float* base, const float* n, const float* wo,
                        uint32_t seed, float* out9) {
    const Mat& m = *reinterpret_cast<const Mat*>(material);
    Rng g;
    g.state = seed;
    Surface s = {};
    s.n = xyz(n);
    s.ng = xyz(n);
    s.front = true;
    BsdfSample b = sample_bsdf(g, m, xyz(base), s, xyz(wo));
    const float r[9] = {b.wi.x, b.wi.y, b.wi.z, b.weight.x, b.weight.y, b.weight.z, b.pdf, b.delta ? 1.0f : 0.0f,
                        b.ok ? 1.0f : 0.0f};
    for (int k = 0; k < 9; ++k) out9[k] = r[k];
}

void oracle_wavelength_rgb(float l, float* rgb) {
    V3 c = wavelength_rgb(l);
    rgb[0] = c.x; rgb[1] = c.y; rgb[2] = c.z;
}

float oracle_env_pdf(const OracleScene* s, const uint8_t* params, const float* d) {
    Ctx C{view(s), reinterpret_cast<const Params*>(params)};
    return env_pdf(C, xyz(d));
}

void oracle_env_radiance(const OracleScene* s, const uint8_t* params, const float* d, float* rgb) {
    Ctx C{view(s), reinterpret_cast<const Params*>(params)};
    V3 c = env_radiance(C, xyz(d));
    rgb[0] = c.x; rgb[1] = c.y; rgb[2] = c.z;
}

}  // extern "C"

// ---------------------------------------------------------------- FP64 denoiser
// The whole denoiser (src/shaders.mjs:1916-2098: prepare, à-trous at steps
// 1, 2, 4, 8, 16, compose) in double precision, from the FP32 sums: the
// error reference for the GPU's FP32 and FP16-guide passes. Same formulas
// and tap order as the FP32 restatement above, every value a double.
namespace {

struct D4 { double x, y, z, w; };

struct Den64 {
    const Params* P;
    const float* accum;
    const float* galb;
    const float* gnorm;
    double samples() const { return (double)std::max(P->total_samples, 1u); }
    static double lum(double r, double g, double b) { return r * 0.2126 + g * 0.7152 + b * 0.0722; }
    D4 albedo(uint32_t i) const {
        const float* g = galb + (size_t)i * 4;
        const double n = samples();
        return {std::max(g[0] / n, 0.01), std::max(g[1] / n, 0.01), std::max(g[2] / n, 0.01), 0.0};
    }
    D4 guide(uint32_t i) const {
        const float* g = gnorm + (size_t)i * 4;
        const double l2 = (double)g[0] * g[0] + (double)g[1] * g[1] + (double)g[2] * g[2];
        const double l = std::sqrt(l2);
        return l2 > 0.0 ? D4{g[0] / l, g[1] / l, g[2] / l, g[3] / samples()} : D4{0.0, 0.0, 0.0, g[3] / samples()};
    }
    static double similarity(const D4& p, const D4& q, double step) {
        if (p.w <= 0.0 || q.w <= 0.0) return (p.w <= 0.0 && q.w <= 0.0) ? 1.0 : 0.0;
        if (p.x * p.x + p.y * p.y + p.z * p.z == 0.0) return 0.0;
        const double wn = std::pow(std::max(p.x * q.x + p.y * q.y + p.z * q.z, 0.0), 64.0);
        const double wz = std::exp(-std::fabs(p.w - q.w) / (p.w * 0.02 * step + 1e-6));
        return wn * wz;
    }
};

const double KERNEL64[3] = {0.375, 0.25, 0.0625};

}  // namespace

extern "C" void oracle_denoise64(const uint8_t* params, const float* accum, const float* guide_albedo,
                                 const float* guide_normal, double* out) {
    Den64 D{reinterpret_cast<const Params*>(params), accum, guide_albedo, guide_normal};
    const Params& P = *D.P;
    const int32_t W = (int32_t)P.width, H = (int32_t)P.height;
    const size_t N = (size_t)W * H;
    const double n = D.samples();
    std::vector<D4> guides(N), light(N), next(N);
    for (size_t i = 0; i < N; ++i) guides[i] = D.guide((uint32_t)i);
    auto light_at = [&](size_t i) {
        const float* a = accum + i * 4;
        const D4 al = D.albedo((uint32_t)i);
        return D4{a[0] / n / al.x, a[1] / n / al.y, a[2] / n / al.z, 0.0};
    };
    for (int32_t y = 0; y < H; ++y)
        for (int32_t x = 0; x < W; ++x) {
            const size_t i = (size_t)y * W + x;
            D4 l = light_at(i);
            const D4 al = D.albedo((uint32_t)i);
            const double a = Den64::lum(al.x, al.y, al.z);
            if (n >= 4.0) {
                const float* ac = accum + i * 4;
                const double mean = Den64::lum(ac[0], ac[1], ac[2]) / n;
                const double spread = std::max(ac[3] / n - mean * mean, 0.0) * n / (n - 1.0);
                l.w = spread / n / (a * a);
            } else {
                double sum = 0.0, sum2 = 0.0, weight = 0.0;
                for (int32_t dy = -2; dy <= 2; dy++)
                    for (int32_t dx = -2; dx <= 2; dx++) {
                        const int32_t xx = x + dx, yy = y + dy;
                        if (xx < 0 || yy < 0 || xx >= W || yy >= H) continue;
                        const size_t j = (size_t)yy * W + xx;
                        const double w = Den64::similarity(guides[i], guides[j], 1.0);
                        const D4 lj = light_at(j);
                        const double v = Den64::lum(lj.x, lj.y, lj.z);
                        sum += w * v;
                        sum2 += w * v * v;
                        weight += w;
                    }
                const double mean = sum / std::max(weight, 1e-6);
                l.w = std::max(sum2 / std::max(weight, 1e-6) - mean * mean, 0.0) + 0.01 * mean * mean;
            }
            light[i] = l;
        }
    for (int32_t step : {1, 2, 4, 8, 16}) {
        for (int32_t y = 0; y < H; ++y)
            for (int32_t x = 0; x < W; ++x) {
                const size_t i = (size_t)y * W + x;
                const D4 c = light[i], gp = guides[i];
                const double lp = Den64::lum(c.x, c.y, c.z);
                double variance = 0.0, vw = 0.0;
                for (int32_t dy = -1; dy <= 1; dy++)
                    for (int32_t dx = -1; dx <= 1; dx++) {
                        const int32_t xx = x + dx, yy = y + dy;
                        if (xx < 0 || yy < 0 || xx >= W || yy >= H) continue;
                        const double w = KERNEL64[std::abs(dx)] * KERNEL64[std::abs(dy)];
                        variance += w * light[(size_t)yy * W + xx].w;
                        vw += w;
                    }
                const double sigma = 4.0 * std::sqrt(variance / vw) + 1e-4;
                double sx = 0.0, sy = 0.0, sz = 0.0, sv = 0.0, weight = 0.0;
                for (int32_t dy = -2; dy <= 2; dy++)
                    for (int32_t dx = -2; dx <= 2; dx++) {
                        const int32_t xx = x + dx * step, yy = y + dy * step;
                        if (xx < 0 || yy < 0 || xx >= W || yy >= H) continue;
                        const size_t j = (size_t)yy * W + xx;
                        const D4 q = light[j];
                        const double w = KERNEL64[std::abs(dx)] * KERNEL64[std::abs(dy)] *
                                         Den64::similarity(gp, guides[j], (double)step) *
                                         std::exp(-std::fabs(lp - Den64::lum(q.x, q.y, q.z)) / sigma);
                        sx += w * q.x;
                        sy += w * q.y;
                        sz += w * q.z;
                        sv += w * w * q.w;
                        weight += w;
                    }
                next[i] = weight > 0.0 ? D4{sx / weight, sy / weight, sz / weight, sv / (weight * weight)} : c;
            }
        std::swap(light, next);
    }
    for (size_t i = 0; i < N; ++i) {
        const D4 al = D.albedo((uint32_t)i);
        out[i * 4] = light[i].x * al.x;
        out[i * 4 + 1] = light[i].y * al.y;
        out[i * 4 + 2] = light[i].z * al.z;
        out[i * 4 + 3] = 0.0;
    }
}
