// Light tracing (k_light_trace), the splat and gather resolve (k_resolve),
// the photon gather (k_photon_gather), and the fold of compensated sums.
#include "path.hpp"

namespace rm {

// ------------------------------------------------------------------ light tracing
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

// Light paths (src/shaders.mjs:1630-1814), photons_n of them a frame: the
// reference's photon_count(P) by default, or a budget the caller sets. A
// block takes photon ids first, first + 1, ... with first advancing by the
// grid's size, so a grid smaller than the budget covers it and the traversal
// spill stays sized by the grid; photon id seeds its own PCG stream, so the
// grid does not change a photon. Deposits happen after the bounce loop with
// one counter atomic per wave (ballot + prefix count): photon slots are
// claimed in a different order than one atomic per lane would give, which
// the reference leaves unspecified too; slots at or past `capacity` are
// dropped, as the reference drops them past its fixed capacity.
template <bool TEXTURES, bool DISPERSION, int TREE>
__global__ __launch_bounds__(BLOCK_THREADS) void k_light_trace(DevScene S, Params P, const float4* __restrict__ E,
                                                               uint32_t* splats, uint32_t* heads,
                                                               float4* __restrict__ photons, uint32_t* spill,
                                                               uint32_t slots, uint32_t photons_n, uint32_t capacity) {
    __shared__ uint2 lds[MAIN_LDS_DEPTH * BLOCK_THREADS];
    const Ctx C{S, &P};
    const uint32_t lane = blockIdx.x * BLOCK_THREADS + threadIdx.x;
    auto stack = BlockStack<MAIN_LDS_DEPTH, TREE>::make(lds, spill, lane, slots);
    // `first` is the same for every thread of a block, so whole waves take
    // each pass and reach the deposit's ballot together.
    for (uint32_t first = blockIdx.x * BLOCK_THREADS; first < photons_n; first += gridDim.x * BLOCK_THREADS) {
        const uint32_t id = first + threadIdx.x;
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
        if (mask == 0) continue;
        uint32_t wave_lane = __lane_id();
        uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
        uint32_t base = 0;
        if (wave_lane == leader) base = atomicAdd(&heads[PHOTON_TABLE], (uint32_t)__popcll(mask));
        base = __shfl(base, (int)leader);
        if (!pending) continue;
        uint32_t i = base + (uint32_t)__popcll(mask & ((1ull << wave_lane) - 1ull));
        if (i >= capacity) continue;
        V3 q = dep_p / P.photon_cell;
        uint32_t h = photon_cell_hash(to_i32(floorf(q.x)), to_i32(floorf(q.y)), to_i32(floorf(q.z)));
        uint32_t next = atomicExch(&heads[h], i + 1u);
        photons[(size_t)i * 3] = make_float4(dep_p.x, dep_p.y, dep_p.z, __uint_as_float(next));
        photons[(size_t)i * 3 + 1] = make_float4(dep_beta.x, dep_beta.y, dep_beta.z, 0.0f);
        photons[(size_t)i * 3 + 2] = make_float4(-dep_d.x, -dep_d.y, -dep_d.z, 0.0f);
    }
}

// resolve_splats() then photon_gather()'s sum into the accumulation
// (src/shaders.mjs:1819-1831, 1610-1624), in that order, for one pixel. The
// gather pass leaves its light in `gathered` (w = 1 where it adds any), so it
// can run while the path-tracing pass still writes the accumulation.
__global__ __launch_bounds__(BLOCK_THREADS) void k_resolve(Params P, uint4* splats, const float4* gathered,
                                                           float4* accum, float4* comp) {
    uint32_t i = blockIdx.x * BLOCK_THREADS + threadIdx.x;
    if (i >= P.width * P.height) return;
    accum[i] = resolve_add(P, splats, gathered, comp, i, accum[i]);
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
        // The hit primitive's material alone decides whether the path ends
        // here (a light, or any surface but a delta one at the first hit),
        // so most camera rays stop before their shading point is built.
        Mat m = load_mat(S, __float_as_uint(S.prims[(size_t)hit.prim * 4].y));
        if (m.kind == LIGHT || (depth == 0u && !is_delta(m))) break;
        Surface s = surface<TEXTURES>(S, o, d, hit);
        V3 base = albedo<TEXTURES>(S, m, s.p, s.uv);
        if (!is_delta(m)) {
            V3 light = throughput * gather(C, heads, photons, s, m, base, -d);
            if (finite3(light)) return make_float4(light.x, light.y, light.z, 1.0f);
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

// ------------------------------------------------------------------ launchers
template <bool T, bool D>
static void launch_light_trace_v(const DevScene& S, const Params& P, const float4* emitters, uint32_t* splats,
                                 uint32_t* heads, float4* photons, uint32_t* spill, uint32_t slots, dim3 grid,
                                 uint32_t count, uint32_t capacity, int tree, hipStream_t stream) {
    with_tree(tree, [&](auto t) {
        hipLaunchKernelGGL((k_light_trace<T, D, decltype(t)::value>), grid, dim3(BLOCK_THREADS), 0, stream, S, P,
                           emitters, splats, heads, photons, spill, slots, count, capacity);
    });
}

void launch_light_trace(const DevScene& S, const Params& P, const float4* emitters, uint32_t* splats, uint32_t* heads,
                        float4* photons, uint32_t* spill, uint32_t slots, uint32_t count, uint32_t capacity,
                        const Features& f, hipStream_t stream) {
    // One photon a thread up to LIGHT_LANES, then each thread takes several.
    const uint32_t lanes = count < LIGHT_LANES ? count : LIGHT_LANES;
    dim3 grid((lanes + BLOCK_THREADS - 1) / BLOCK_THREADS);
    if (grid.x == 0) return;
    if (f.textures)
        launch_light_trace_v<true, true>(S, P, emitters, splats, heads, photons, spill, slots, grid, count, capacity,
                                         f.tree, stream);
    else if (f.dispersion)
        launch_light_trace_v<false, true>(S, P, emitters, splats, heads, photons, spill, slots, grid, count,
                                          capacity, f.tree, stream);
    else
        launch_light_trace_v<false, false>(S, P, emitters, splats, heads, photons, spill, slots, grid, count,
                                           capacity, f.tree, stream);
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

}  // namespace rm
