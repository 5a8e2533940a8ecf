// Device-side shading for the Raymond kernels: surfaces, textures, the sky
// and environment map, materials (diffuse, GGX metal and plastic, exact
// Fresnel glass, Cauchy dispersion), light sampling with MIS, homogeneous
// volumes with Henyey-Greenstein phase, and the path integrator. Each
// function follows the reference WGSL cited beside it (src/shaders.mjs).
#pragma once
#include "device.hpp"

namespace rm {

// ------------------------------------------------------------------ textures
// Bilinear with repeat addressing (sample() in src/Texture.mjs:164-187, the
// reference's filtering sampler); sRGB bytes decode before blending.
RM_D float srgb_decode(uint32_t b) {
    float c = (float)b / 255.0f;
    return c <= 0.04045f ? c / 12.92f : powf((c + 0.055f) / 1.055f, 2.4f);
}
RM_D V3 tex_sample(const uchar4* data, uint32_t w, uint32_t h, uint32_t layer, V2 uv, bool srgb) {
    float x = uv.x * (float)w - 0.5f, y = uv.y * (float)h - 0.5f;
    float fx0 = floorf(x), fy0 = floorf(y);
    float fx = x - fx0, fy = y - fy0;
    int64_t x0 = (int64_t)fx0, y0 = (int64_t)fy0;
    const float ws[4] = {(1.0f - fx) * (1.0f - fy), fx * (1.0f - fy), (1.0f - fx) * fy, fx * fy};
    V3 acc = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        int64_t px = ((x0 + (k & 1)) % (int64_t)w + w) % w;
        int64_t py = ((y0 + (k >> 1)) % (int64_t)h + h) % h;
        uchar4 t = data[(size_t)layer * w * h + (size_t)py * w + px];
        V3 c = srgb ? v3(srgb_decode(t.x), srgb_decode(t.y), srgb_decode(t.z))
                    : v3(t.x / 255.0f, t.y / 255.0f, t.z / 255.0f);
        acc = acc + c * ws[k];
    }
    return acc;
}

// ------------------------------------------------------------------ surface
struct Surface { V3 p, n, ng; bool front; uint32_t material, light, kind; V2 uv; };

RM_D V2 sphere_uv(V3 n, float4 p1, float4 p2) {
    V3 z = cross(xyz(p1), xyz(p2));
    float u = atan2f(-dot(n, z), dot(n, xyz(p1))) / TWO_PI;
    return {u - floorf(u), acosf(clampf(dot(n, xyz(p2)), -1.0f, 1.0f)) / PI};
}

RM_D V3 bump(V3 n, V3 ng, V3 tangent, V3 up, V3 t, float strength) {
    V3 right = tangent - n * dot(n, tangent);
    if (length(right) < 1e-9f) return n;
    right = normalize(right);
    V3 above = cross(n, right);
    if (dot(above, up) < 0.0f) above = -above;
    V3 tilted = normalize(right * ((t.x * 2.0f - 1.0f) * strength) + above * ((t.y * 2.0f - 1.0f) * strength) +
                          n * (t.z * 2.0f - 1.0f));
    return dot(tilted, ng) > 0.0f ? tilted : n;
}

// Hit → shading point, normals facing the ray, uv (src/shaders.mjs:460-569).
template <bool TEXTURES>
RM_D Surface surface(const DevScene& S, V3 o, V3 d, const Hit& hit) {
    const float4* pr = S.prims + (size_t)hit.prim * 4;
    float4 hd = pr[0], p0 = pr[1], p1 = pr[2], p2 = pr[3];
    uint32_t kind = __float_as_uint(hd.x), material = __float_as_uint(hd.y), light = __float_as_uint(hd.z),
             attributes = __float_as_uint(hd.w);
    V3 point = o + d * hit.t;
    V3 ng, ns, tangent = {0, 0, 0}, up = {0, 0, 0};
    V2 uv;
    if (kind == SPHERE) {
        ng = normalize(point - xyz(p0));
        ns = ng;
        uv = sphere_uv(ng, p1, p2);
        if (TEXTURES) {
            tangent = cross(xyz(p2), ng);
            if (length(tangent) < 1e-6f) tangent = xyz(p1);
            up = cross(ng, tangent);
        }
    } else if (kind == QUAD) {
        ng = normalize(cross(xyz(p1), xyz(p2)));
        ns = ng;
        uv = {hit.u, 1.0f - hit.v};
        tangent = xyz(p1);
        up = xyz(p2);
    } else if (kind == BOX) {
        V3 lp = {dot(xyz(p0), point) + p0.w, dot(xyz(p1), point) + p1.w, dot(xyz(p2), point) + p2.w};
        V3 a = {fabsf(lp.x), fabsf(lp.y), fabsf(lp.z)};
        V3 half = {1.0f / length(xyz(p0)), 1.0f / length(xyz(p1)), 1.0f / length(xyz(p2))};
        if (a.x >= a.y && a.x >= a.z) {
            ng = xyz(p0) * signf(lp.x);
            uv = {(1.0f - signf(lp.x) * lp.z) * half.z, (1.0f - lp.y) * half.y};
            tangent = (-signf(lp.x)) * xyz(p2);
            up = xyz(p1);
        } else if (a.y >= a.z) {
            ng = xyz(p1) * signf(lp.y);
            uv = {(1.0f + lp.x) * half.x, (1.0f + signf(lp.y) * lp.z) * half.z};
            tangent = xyz(p0);
            up = (-signf(lp.y)) * xyz(p2);
        } else {
            ng = xyz(p2) * signf(lp.z);
            uv = {(1.0f + signf(lp.z) * lp.x) * half.x, (1.0f - lp.y) * half.y};
            tangent = signf(lp.z) * xyz(p0);
            up = xyz(p1);
        }
        ng = normalize(ng);
        ns = ng;
    } else {
        V3 e1 = xyz(p1) - xyz(p0), e2 = xyz(p2) - xyz(p0);
        ng = normalize(cross(e1, e2));
        ns = ng;
        uv = {hit.u, hit.v};
        V2 du1 = {1.0f, 0.0f}, du2 = {0.0f, 1.0f};
        if (attributes != NONE) {
            float w = 1.0f - hit.u - hit.v;
            const float4* A = S.attr + attributes;
            float4 a0 = A[0], a1 = A[1], a2 = A[2], vs = A[3];
            uint32_t flags = to_u32(vs.w);
            if (flags & 1u) {
                ns = normalize(xyz(a0) * w + xyz(a1) * hit.u + xyz(a2) * hit.v);
                if (dot(ns, ng) < 0.0f) ns = -ns;
            }
            if (flags & 2u) {
                uv = {a0.w * w + a1.w * hit.u + a2.w * hit.v, vs.x * w + vs.y * hit.u + vs.z * hit.v};
                V2 t1 = {a1.w - a0.w, vs.y - vs.x};
                V2 t2 = {a2.w - a0.w, vs.z - vs.x};
                if (fabsf(t1.x * t2.y - t2.x * t1.y) > 1e-12f) { du1 = t1; du2 = t2; }
            }
        }
        if (TEXTURES) {
            float det = du1.x * du2.y - du2.x * du1.y;
            tangent = (e1 * du2.y - e2 * du1.y) / det;
            up = -((e2 * du1.x - e1 * du2.x) / det);
        }
    }
    if (TEXTURES) {
        const float4* m = S.mats + (size_t)material * 6;
        uint32_t nmap = __float_as_uint(m[4].x);
        if (nmap != NONE && S.nmap_layers > 0) {
            float4 ts = m[3];
            V3 t = tex_sample(S.nmaps, S.nmap_w, S.nmap_h, nmap, {uv.x * ts.z, uv.y * ts.w}, false);
            ns = bump(ns, ng, tangent, up, t, m[4].y);
        }
    }
    bool front = dot(ng, d) < 0.0f;
    float flip = front ? 1.0f : -1.0f;
    return {point, ns * flip, ng * flip, front, material, light, kind, uv};
}

RM_D V3 offset(const Surface& s, V3 wi) {
    float eps = 1e-4f * fmaxf(1.0f, fmaxf(fabsf(s.p.x), fmaxf(fabsf(s.p.y), fabsf(s.p.z))));
    return s.p + s.ng * (dot(wi, s.ng) > 0.0f ? eps : -eps);
}

// ------------------------------------------------------------------ samplers
struct Frame { V3 c0, c1, c2; };
RM_D Frame frame(V3 n) {
    float s = n.z >= 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (s + n.z);
    float b = n.x * n.y * a;
    return {{1.0f + s * n.x * n.x * a, s * b, -s * n.x}, {b, s + n.y * n.y * a, -n.y}, n};
}
RM_D V3 fmul(const Frame& f, V3 l) { return f.c0 * l.x + f.c1 * l.y + f.c2 * l.z; }

RM_D V3 cosine_hemisphere(Rng& g, V3 n) {
    V2 u = g.rand2();
    float r = sqrtf(u.x);
    float phi = TWO_PI * u.y;
    return normalize(fmul(frame(n), {r * cosf(phi), r * sinf(phi), sqrtf(fmaxf(0.0f, 1.0f - r * r))}));
}
RM_D V3 sample_cone(Rng& g, V3 axis, float omc) {
    V2 u = g.rand2();
    float cos_t = 1.0f - u.x * omc;
    float sin_t = sqrtf(fmaxf(0.0f, 1.0f - cos_t * cos_t));
    float phi = TWO_PI * u.y;
    return normalize(fmul(frame(axis), {sin_t * cosf(phi), sin_t * sinf(phi), cos_t}));
}
RM_D V2 concentric_disk(Rng& g) {
    V2 r = g.rand2();
    V2 u = {2.0f * r.x - 1.0f, 2.0f * r.y - 1.0f};
    if (u.x == 0.0f && u.y == 0.0f) return {0.0f, 0.0f};
    if (fabsf(u.x) > fabsf(u.y)) {
        float phi = PI_4 * (u.y / u.x);
        return {u.x * cosf(phi), u.x * sinf(phi)};
    }
    float phi = PI_2 - PI_4 * (u.x / u.y);
    return {u.y * cosf(phi), u.y * sinf(phi)};
}

// ------------------------------------------------------------------ sky
RM_D float half_bits(unsigned short h) { return (float)__builtin_bit_cast(_Float16, h); }

struct Ctx {
    DevScene S;
    const Params* P;   // kernel-argument copy, read with scalar loads
};

RM_D V2 env_uv(const Params& P, V3 d) {
    float u = 0.5f + atan2f(d.x, -d.z) / TWO_PI - P.env_rotation / TWO_PI;
    return {u - floorf(u), acosf(clampf(d.y, -1.0f, 1.0f)) / PI};
}
RM_D V3 env_radiance(const Ctx& C, V3 d) {
    const DevScene& S = C.S;
    V2 uv = env_uv(*C.P, d);
    uint32_t w = S.env_w, h = S.env_h;
    float x = uv.x * (float)w - 0.5f, y = uv.y * (float)h - 0.5f;
    float fx0 = floorf(x), fy0 = floorf(y);
    float fx = x - fx0, fy = y - fy0;
    int64_t x0 = (int64_t)fx0, y0 = (int64_t)fy0;
    const float ws[4] = {(1.0f - fx) * (1.0f - fy), fx * (1.0f - fy), (1.0f - fx) * fy, fx * fy};
    V3 acc = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        int64_t px = ((x0 + (k & 1)) % (int64_t)w + w) % w;
        int64_t py = y0 + (k >> 1);
        py = py < 0 ? 0 : (py > (int64_t)h - 1 ? (int64_t)h - 1 : py);
        ushort4 t = S.env[(size_t)py * w + px];
        acc = acc + v3(half_bits(t.x), half_bits(t.y), half_bits(t.z)) * ws[k];
    }
    return acc * C.P->env_strength;
}
RM_D float env_cdf(const DevScene& S, uint32_t x, uint32_t y) { return S.env_table[(size_t)y * S.table_w + x]; }
RM_D uint32_t env_search(const DevScene& S, uint32_t y, uint32_t n, float xi) {
    uint32_t lo = 0u, hi = n - 1u;
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2u;
        if (env_cdf(S, mid, y) > xi) hi = mid;
        else lo = mid + 1u;
    }
    return lo;
}
RM_D float env_pdf(const Ctx& C, V3 d) {
    const DevScene& S = C.S;
    uint32_t cw = S.table_w, ch = S.table_h;
    V2 uv = env_uv(*C.P, d);
    uint32_t x = min(to_u32(uv.x * (float)cw), cw - 1u);
    uint32_t y = min(to_u32(uv.y * (float)ch), ch - 1u);
    float row = env_cdf(S, y, ch) - (y > 0u ? env_cdf(S, y - 1u, ch) : 0.0f);
    float cell = env_cdf(S, x, y) - (x > 0u ? env_cdf(S, x - 1u, y) : 0.0f);
    float s = sinf(uv.y * PI);
    return s > 0.0f ? row * cell * (float)(cw * ch) / (TWO_PI_PI * s) : 0.0f;
}
RM_D V3 sample_env(const Ctx& C, Rng& g) {
    const DevScene& S = C.S;
    uint32_t cw = S.table_w, ch = S.table_h;
    V2 cell = g.rand2();
    uint32_t y = env_search(S, ch, ch, cell.y);
    uint32_t x = env_search(S, y, cw, cell.x);
    V2 within = g.rand2();
    float u = ((float)x + within.x) / (float)cw + C.P->env_rotation / TWO_PI;
    float v = ((float)y + within.y) / (float)ch;
    float phi = (u - 0.5f) * 2.0f * PI;
    float theta = v * PI;
    return {sinf(theta) * sinf(phi), cosf(theta), -sinf(theta) * cosf(phi)};
}
template <bool ENV>
RM_D bool has_env(const Ctx& C) { return ENV && C.P->env_strength > 0.0f; }

template <bool ENV>
RM_D V3 sky(const Ctx& C, V3 d) {
    const Params& P = *C.P;
    if (has_env<ENV>(C)) return env_radiance(C, d);
    if (d.y >= 0.0f) return mix(pv3(P.sky_horizon), pv3(P.sky_zenith), sqrtf(fminf(d.y, 1.0f)));
    return mix(pv3(P.sky_horizon), pv3(P.sky_ground), sqrtf(fminf(-d.y, 1.0f)));
}
RM_D bool in_sun(const Params& P, V3 d) { return P.sun_on != 0u && dot(d, pv3(P.sun_dir)) >= 1.0f - P.sun_omc; }

// ------------------------------------------------------------------ materials
template <bool TEXTURES>
RM_D V3 albedo(const DevScene& S, const Mat& m, V3 p, V2 uv) {
    // Component-wise selects: choosing between two struct members as a whole
    // makes the compiler index the struct in private memory.
    bool odd = false;
    if (m.checker_size > 0.0f) {
        V3 q = p / m.checker_size + s3(1e-3f);
        int32_t cx = to_i32(floorf(q.x)), cy = to_i32(floorf(q.y)), cz = to_i32(floorf(q.z));
        odd = (((uint32_t)cx + (uint32_t)cy + (uint32_t)cz) & 1u) != 0u;
    }
    V3 c = {odd ? m.checker.x : m.color.x, odd ? m.checker.y : m.color.y, odd ? m.checker.z : m.color.z};
    if (TEXTURES && m.texture != NONE && S.tex_layers > 0)
        c = c * tex_sample(S.textures, S.tex_w, S.tex_h, m.texture, {uv.x * m.tsx, uv.y * m.tsy}, true);
    return c;
}
template <bool TEXTURES>
RM_D V3 emitted(const DevScene& S, const Mat& m, V2 uv) {
    if (!TEXTURES || m.emission_texture == NONE || S.tex_layers == 0) return m.emission;
    return m.emission * tex_sample(S.textures, S.tex_w, S.tex_h, m.emission_texture, {uv.x * m.tsx, uv.y * m.tsy}, true);
}
RM_D bool is_delta(const Mat& m) { return m.kind == GLASS || (m.kind == METAL && m.roughness < 0.01f); }
RM_D float ggx_alpha2(float r) { float a = fmaxf(r * r, 0.002f); return a * a; }
RM_D float ggx_d(float nh, float a2) { float k = nh * nh * (a2 - 1.0f) + 1.0f; return a2 / (PI * k * k); }
RM_D float smith_g1(float nv, float a2) { return 2.0f * nv / (nv + sqrtf(a2 + (1.0f - a2) * nv * nv)); }
RM_D V3 sample_ggx(Rng& g, V3 n, float a2) {
    V2 u = g.rand2();
    float r = u.x;
    float cos_t = sqrtf((1.0f - r) / (1.0f + (a2 - 1.0f) * r));
    float sin_t = sqrtf(fmaxf(0.0f, 1.0f - cos_t * cos_t));
    float phi = TWO_PI * u.y;
    return normalize(fmul(frame(n), {sin_t * cosf(phi), sin_t * sinf(phi), cos_t}));
}
RM_D float pow5(float c) { return powf(1.0f - clampf(c, 0.0f, 1.0f), 5.0f); }
RM_D V3 schlick(V3 f0, float c) { float k = pow5(c); return f0 + (s3(1.0f) - f0) * k; }
RM_D float schlick1(float f0, float c) { float k = pow5(c); return f0 + (1.0f - f0) * k; }
RM_D float dielectric_f0(float ior) { float r = (ior - 1.0f) / (ior + 1.0f); return r * r; }
RM_D float fresnel_dielectric(float cos_i, float eta) {
    float sin2_t = (1.0f - cos_i * cos_i) / (eta * eta);
    if (sin2_t >= 1.0f) return 1.0f;
    float cos_t = sqrtf(1.0f - sin2_t);
    float rs = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
    float rp = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
    return 0.5f * (rs * rs + rp * rp);
}
RM_D float plastic_spec_chance(const Mat& m, float nv) {
    return clampf(schlick1(dielectric_f0(m.ior), nv) * 4.0f + 0.1f, 0.1f, 0.9f);
}

struct Eval { V3 f; float pdf; };
RM_D Eval eval_bsdf(const Mat& m, V3 base, V3 n, V3 wo, V3 wi) {
    float nl = dot(n, wi), nv = dot(n, wo);
    if (nl <= 0.0f || nv <= 0.0f) return {{0.0f, 0.0f, 0.0f}, 0.0f};
    if (m.kind == DIFFUSE) return {base * (nl / PI), nl / PI};
    float a2 = ggx_alpha2(m.roughness);
    V3 h = normalize(wo + wi);
    float nh = fmaxf(dot(n, h), 0.0f);
    float vh = fmaxf(dot(wo, h), 1e-6f);
    float d = ggx_d(nh, a2);
    float g = smith_g1(nv, a2) * smith_g1(nl, a2);
    float spec_pdf = d * nh / (4.0f * vh);
    if (m.kind == METAL) return {schlick(base, vh) * (d * g / (4.0f * nv)), spec_pdf};
    float f0 = dielectric_f0(m.ior);
    float coat = schlick1(f0, vh) * d * g / (4.0f * nv);
    float through = (1.0f - schlick1(f0, nl)) * (1.0f - schlick1(f0, nv));
    V3 diffuse = base * (through * nl / PI);
    float chance = plastic_spec_chance(m, nv);
    return {s3(coat) + diffuse, chance * spec_pdf + (1.0f - chance) * nl / PI};
}

struct BsdfSample { V3 wi, weight; float pdf; bool delta, ok; };
RM_D BsdfSample sample_bsdf(Rng& g, const Mat& m, V3 base, const Surface& s, V3 wo) {
    V3 n = s.n;
    float nv = dot(n, wo);
    if (m.kind == GLASS) {
        float eta = s.front ? m.ior : 1.0f / m.ior;
        float cos_i = clampf(nv, 0.0f, 1.0f);
        float f = fresnel_dielectric(cos_i, eta);
        if (g.rand() < f) return {reflect(-wo, n), {1, 1, 1}, 0.0f, true, true};
        V3 wi = refract(-wo, n, 1.0f / eta);
        if (dot(wi, wi) == 0.0f) return {reflect(-wo, n), {1, 1, 1}, 0.0f, true, true};
        return {normalize(wi), s.front ? base : v3(1, 1, 1), 0.0f, true, true};
    }
    if (m.kind == METAL && m.roughness < 0.01f) return {reflect(-wo, n), schlick(base, nv), 0.0f, true, true};
    V3 wi;
    if (m.kind == DIFFUSE) {
        wi = cosine_hemisphere(g, n);
    } else if (m.kind == METAL || g.rand() < plastic_spec_chance(m, nv)) {
        wi = reflect(-wo, sample_ggx(g, n, ggx_alpha2(m.roughness)));
    } else {
        wi = cosine_hemisphere(g, n);
    }
    const BsdfSample fail = {{0, 0, 0}, {0, 0, 0}, 0.0f, false, false};
    if (dot(wi, s.ng) <= 0.0f) return fail;
    Eval e = eval_bsdf(m, base, n, wo, wi);
    if (e.pdf <= 0.0f) return fail;
    return {wi, e.f / e.pdf, e.pdf, false, true};
}

// ------------------------------------------------------------------ lights
template <bool ENV>
RM_D uint32_t light_choices(const Ctx& C) {
    return C.P->light_count + (C.P->sun_on != 0u ? 1u : 0u) + (has_env<ENV>(C) ? 1u : 0u);
}
RM_D float light_cdf(const Ctx& C, uint32_t l) {
    return __uint_as_float(C.S.lights[C.P->light_count + C.P->volume_count + l]);
}
template <bool ENV>
RM_D float light_chance(const Ctx& C, uint32_t l) {
    float below = l > 0u ? light_cdf(C, l - 1u) : 0.0f;
    return (light_cdf(C, l) - below) * (float)C.P->light_count / (float)light_choices<ENV>(C);
}
RM_D uint32_t pick_light(const Ctx& C, float u) {
    uint32_t lo = 0u, hi = C.P->light_count - 1u;
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2u;
        if (light_cdf(C, mid) > u) hi = mid;
        else lo = mid + 1u;
    }
    return lo;
}
RM_D float sphere_omc(float r, float d2) {
    float x = fminf(r * r / d2, 1.0f);
    return x / (1.0f + sqrtf(1.0f - x));
}

struct LightSample { V3 wi; float dist; V3 radiance; float pdf; };

// One of {lights, sun, environment}, then a point or direction on it
// (src/shaders.mjs:931-990).
template <bool ENV, bool TEXTURES>
RM_D LightSample sample_light(const Ctx& C, Rng& g, V3 p) {
    const Params& P = *C.P;
    const LightSample none = {{0, 0, 0}, 0.0f, {0, 0, 0}, 0.0f};
    uint32_t choices = light_choices<ENV>(C);
    float x = g.rand() * (float)choices;
    uint32_t pick = min(to_u32(x), choices - 1u);
    float chance = 1.0f / (float)choices;
    if (has_env<ENV>(C) && pick == P.light_count + (P.sun_on != 0u ? 1u : 0u)) {
        V3 wi = sample_env(C, g);
        float pdf = env_pdf(C, wi);
        if (pdf <= 0.0f) return none;
        return {wi, INF, env_radiance(C, wi), chance * pdf};
    }
    if (P.sun_on != 0u && pick == P.light_count) {
        V3 wi = sample_cone(g, pv3(P.sun_dir), P.sun_omc);
        return {wi, INF, pv3(P.sun_radiance), chance / (TWO_PI * P.sun_omc)};
    }
    pick = pick_light(C, x / (float)P.light_count);
    chance = light_chance<ENV>(C, pick);
    const float4* pr = C.S.prims + (size_t)C.S.lights[pick] * 4;
    float4 hd = pr[0], p0 = pr[1], p1 = pr[2], p2 = pr[3];
    Mat m = load_mat(C.S, __float_as_uint(hd.y));
    if (__float_as_uint(hd.x) == QUAD) {
        V2 ab = g.rand2();
        float a = ab.x, b = ab.y;
        V3 q = xyz(p0) + xyz(p1) * a + xyz(p2) * b;
        V3 to = q - p;
        float d2 = dot(to, to);
        float dist = sqrtf(d2);
        V3 wi = to / dist;
        V3 c = cross(xyz(p1), xyz(p2));
        float area = length(c);
        float cos_l = -dot(wi, c / area);
        if (cos_l <= 0.0f) return none;
        return {wi, dist, emitted<TEXTURES>(C.S, m, {a, 1.0f - b}), chance * d2 / (cos_l * area)};
    }
    V3 to = xyz(p0) - p;
    float d2 = dot(to, to);
    float r = p0.w;
    if (d2 <= r * r) return none;
    float omc = sphere_omc(r, d2);
    V3 wi = sample_cone(g, to / sqrtf(d2), omc);
    float b = dot(to, wi);
    float dist = b - sqrtf(fmaxf(b * b - (d2 - r * r), 0.0f));
    V3 n = normalize(p + wi * dist - xyz(p0));
    V2 uv = TEXTURES ? sphere_uv(n, p1, p2) : V2{0.0f, 0.0f};
    return {wi, dist, emitted<TEXTURES>(C.S, m, uv), chance / (TWO_PI * omc)};
}

template <bool ENV>
RM_D float light_pdf(const Ctx& C, uint32_t prim_index, V3 p, V3 wi, float t) {
    const float4* pr = C.S.prims + (size_t)prim_index * 4;
    float4 hd = pr[0], p0 = pr[1];
    float chance = light_chance<ENV>(C, __float_as_uint(hd.z));
    if (__float_as_uint(hd.x) == QUAD) {
        V3 c = cross(xyz(pr[2]), xyz(pr[3]));
        float area = length(c);
        float cos_l = fabsf(dot(wi, c / area));
        return chance * t * t / fmaxf(cos_l * area, 1e-12f);
    }
    V3 to = xyz(p0) - p;
    return chance / (TWO_PI * sphere_omc(p0.w, dot(to, to)));
}

RM_D float power_heuristic(float a, float b) {
    float a2 = a * a, b2 = b * b;
    if (a2 + b2 <= 0.0f) return 0.0f;
    return a2 / (a2 + b2);
}

// ------------------------------------------------------------------ volumes
RM_D V2 volume_span(const DevScene& S, uint32_t i, V3 o, V3 d) {
    const float4* pr = S.prims + (size_t)i * 4;
    float4 p0 = pr[1];
    if (__float_as_uint(pr[0].x) == SPHERE) {
        V3 oc = o - xyz(p0);
        float a = dot(d, d);
        float half_b = dot(oc, d);
        float disc = half_b * half_b - a * (dot(oc, oc) - p0.w * p0.w);
        if (disc < 0.0f) return {1.0f, 0.0f};
        float sq = sqrtf(disc);
        return {fmaxf((-half_b - sq) / a, 0.0f), (-half_b + sq) / a};
    }
    float4 p1 = pr[2], p2 = pr[3];
    V3 lo = {dot(xyz(p0), o) + p0.w, dot(xyz(p1), o) + p1.w, dot(xyz(p2), o) + p2.w};
    V3 sd = safe({dot(xyz(p0), d), dot(xyz(p1), d), dot(xyz(p2), d)});
    V3 inv = {1.0f / sd.x, 1.0f / sd.y, 1.0f / sd.z};
    V3 t0 = (s3(-1.0f) - lo) * inv, t1 = (s3(1.0f) - lo) * inv;
    V3 tmin = vmin(t0, t1), tmax = vmax(t0, t1);
    return {fmaxf(fmaxf(fmaxf(tmin.x, tmin.y), tmin.z), 0.0f), fminf(fminf(tmax.x, tmax.y), tmax.z)};
}
RM_D uint32_t volume_prim(const Ctx& C, uint32_t k) { return C.S.lights[C.P->light_count + k]; }
RM_D float volume_density(const DevScene& S, uint32_t i) {
    return S.mats[(size_t)__float_as_uint(S.prims[(size_t)i * 4].y) * 6 + 5].x;
}

struct Scatter { float t; uint32_t volume; };
RM_D Scatter scatter_in_volumes(const Ctx& C, Rng& g, V3 o, V3 d, float t_max) {
    Scatter s = {t_max, NONE};
    for (uint32_t k = 0; k < C.P->volume_count; k++) {
        uint32_t i = volume_prim(C, k);
        V2 span = volume_span(C.S, i, o, d);
        float end = fminf(span.y, s.t);
        if (span.x < end) {
            float t = span.x - logf(1.0f - g.rand()) / volume_density(C.S, i);
            if (t < end) s = {t, i};
        }
    }
    return s;
}
RM_D float transmittance(const Ctx& C, V3 o, V3 d, float dist) {
    float depth = 0.0f;
    for (uint32_t k = 0; k < C.P->volume_count; k++) {
        uint32_t i = volume_prim(C, k);
        V2 span = volume_span(C.S, i, o, d);
        float end = fminf(span.y, dist);
        if (span.x < end) depth += volume_density(C.S, i) * (end - span.x);
    }
    return expf(-depth);
}
RM_D float henyey_greenstein(float c, float g) {
    float k = 1.0f + g * g - 2.0f * g * c;
    return (1.0f - g * g) / (FOUR_PI * k * sqrtf(k));
}
RM_D V3 sample_henyey_greenstein(Rng& rg, V3 d, float g) {
    V2 u = rg.rand2();
    float c = 1.0f - 2.0f * u.x;
    if (fabsf(g) >= 1e-3f) {
        float sq = (1.0f - g * g) / (1.0f + g - 2.0f * g * u.x);
        c = clampf((1.0f + g * g - sq * sq) / (2.0f * g), -1.0f, 1.0f);
    }
    float sin_t = sqrtf(fmaxf(0.0f, 1.0f - c * c));
    float phi = TWO_PI * u.y;
    return normalize(fmul(frame(d), {sin_t * cosf(phi), sin_t * sinf(phi), c}));
}

// ------------------------------------------------------------------ dispersion
RM_D float lobe(float x, float mu, float below, float above) {
    float t = (x - mu) / (x < mu ? below : above);
    return expf(-0.5f * t * t);
}
RM_D V3 wavelength_rgb(float l) {
    float x = 1.056f * lobe(l, 599.8f, 37.9f, 31.0f) + 0.362f * lobe(l, 442.0f, 16.0f, 26.7f) -
              0.065f * lobe(l, 501.1f, 20.4f, 26.2f);
    float y = 0.821f * lobe(l, 568.8f, 46.9f, 40.5f) + 0.286f * lobe(l, 530.9f, 16.3f, 31.1f);
    float z = 1.217f * lobe(l, 437.0f, 11.8f, 36.0f) + 0.681f * lobe(l, 459.0f, 26.0f, 13.8f);
    V3 rgb = {3.2406f * x - 1.5372f * y - 0.4986f * z, -0.9689f * x + 1.8758f * y + 0.0415f * z,
              0.0557f * x - 0.2040f * y + 1.0570f * z};
    return vmax(rgb, s3(0.0f)) * v3(1.92985f, 2.94689f, 3.11081f);
}
// The index of refraction a path sees at glass: its own, or at the path's
// wavelength for glass that splits colours, choosing the wavelength (and
// weighting the throughput by its colour) at the first such glass.
template <bool DISPERSION>
RM_D float dispersed_ior(const Mat& m, float& wavelength, V3& throughput, float u) {
    if (DISPERSION && m.kind == GLASS && m.dispersion > 0.0f) {
        if (wavelength == 0.0f) {
            wavelength = 380.0f + 340.0f * u;
            throughput = throughput * wavelength_rgb(wavelength);
        }
        float l = wavelength * 1e-3f;
        return m.ior + m.dispersion * (1.0f / (l * l) - INV_D_LINE2);
    }
    return m.ior;
}

// ------------------------------------------------------------------ camera
struct Ray { V3 o, d; };
RM_D Ray camera_ray(const Params& P, Rng& g, V2 px, bool jitter) {
    V2 ndc = {2.0f * px.x / (float)P.width - 1.0f, 1.0f - 2.0f * px.y / (float)P.height};
    V3 dir = normalize(pv3(P.cam_forward) + pv3(P.cam_right) * (ndc.x * P.tan_half_fov * P.aspect) +
                       pv3(P.cam_up) * (ndc.y * P.tan_half_fov));
    if (!jitter || P.lens_radius <= 0.0f) return {pv3(P.cam_pos), dir};
    V3 focus = pv3(P.cam_pos) + dir * (P.focus_distance / dot(dir, pv3(P.cam_forward)));
    V2 lens = concentric_disk(g);
    lens = {lens.x * P.lens_radius, lens.y * P.lens_radius};
    V3 origin = pv3(P.cam_pos) + pv3(P.cam_right) * lens.x + pv3(P.cam_up) * lens.y;
    return {origin, normalize(focus - origin)};
}

RM_D bool roulette(Rng& g, V3& throughput, uint32_t depth) {
    if (depth < 3u) return true;
    V3 t = throughput;
    float survive = fminf(fmaxf(t.x, fmaxf(t.y, t.z)), 0.95f);
    if (g.rand() >= survive) return false;
    throughput = t / survive;
    return true;
}

}  // namespace rm
