// Light tracing (k_light_trace), the photon map's counting sort
// (k_photon_scan, k_photon_sort), the splat and gather
// resolve (k_resolve), the photon gather (k_gather_begin, k_gather_paths,
// k_gather_walk), and the fold of compensated sums.
#include "path.hpp"

namespace rm {

// ------------------------------------------------------------------ light tracing
RM_D uint32_t photon_cell_hash(int32_t cx, int32_t cy, int32_t cz) {
    return (((uint32_t)cx * 73856093u) ^ ((uint32_t)cy * 19349663u) ^ ((uint32_t)cz * 83492791u)) & (PHOTON_TABLE - 1u);
}

// Whether the ray's slab test passes a box of the scene's delta cut: the
// same reciprocal direction and slab test as the FP32 walk's, so a ray that
// passes none is one whose walk tests no delta primitive.
RM_D bool may_meet_delta(const DevScene& S, V3 o, V3 d) {
    const V3 sd = safe(d);
    const V3 inv = {1.0f / sd.x, 1.0f / sd.y, 1.0f / sd.z};
    for (uint32_t i = 0u; i < S.delta_count; ++i)
        if (slabs(S.delta_boxes[2 * i], S.delta_boxes[2 * i + 1], o, inv, INF) < INF) return true;
    return false;
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

// The light-tracing and gather kernels run beside the path pass and are
// compiled for more waves per SIMD than it (side_waves): five (96 VGPRs)
// beside a path pass at four, seven (72 VGPRs) beside one at five, so a wave
// of theirs fits where a path-pass wave retires. At more than 96, on the
// caustics scene, light tracing waited for two to retire and took 2.5 ms a
// frame beside the path pass instead of 0.64; at seven beside a path pass at
// five, the 8 M-photon caustics frame ran 5% faster than at five, and at six
// beside a path pass at four, the sphere frame 0.5% slower.
// k_light_trace_batch, beside the batched path pass (four waves), stays at
// five.
//
// Light paths (src/shaders.mjs:1630-1814), photons_n of them a frame: the
// reference's photon_count(P) by default, or a budget the caller sets. A
// block takes photon ids first, first + 1, ... with first advancing by the
// grid's size, so a grid smaller than the budget covers it and the traversal
// spill stays sized by the grid; photon id seeds its own PCG stream, so the
// grid does not change a photon. Deposits happen after the bounce loop with
// one counter atomic per wave (ballot + prefix count): photon slots are
// claimed in a different order than one atomic per lane would give, which
// the reference leaves unspecified too; slots at or past `capacity` are
// dropped, as the reference drops them past its fixed capacity. A deposit
// counts its photon in its cell (the rank it gets orders the cell's photons
// for the sort) where the reference links it into the cell's list.
//
// light_paths is one block's share, block `block` of a grid of `blocks`;
// k_light_trace runs a frame's grid, k_light_trace_batch several jobs' grids
// in one launch.
template <bool TEXTURES, bool DISPERSION, class STACK>
RM_D void light_paths(const DevScene& S, const Params& P, const float4* __restrict__ E, uint32_t* splats,
                      const PhotonMap& M, STACK& stack, uint32_t photons_n, uint32_t block, uint32_t blocks) {
    constexpr int TREE = STACK::tree;
    const Ctx C{S, &P};
    // `first` is the same for every thread of a block, so whole waves take
    // each pass and reach the deposit's ballot together.
    for (uint32_t first = block * BLOCK_THREADS; first < photons_n; first += blocks * BLOCK_THREADS) {
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
                    // How many targets' cones hold d. Exactly, d is in the
                    // cone of a target at `to` (|to|² = d2 > r², the origin
                    // outside) iff s = d·to >= 0 and s² >= d2 - r². A target
                    // is skipped (the test adds 0) where that fails by a
                    // margin far beyond the test's rounding, a few ulps of
                    // the cosines: s < 0 with s² > 1e-8 d2 (the computed
                    // cosine is negative, so 1 - cos > 1 >= omc), or s >= 0
                    // with s² < 0.999 (d2 - r²) - 1e-4 d2 (cos θ < cos α -
                    // 3e-4). Only near-misses take the divisions.
                    float within = 1.0f;
                    for (uint32_t i = 0u; i < targets; i++) {
                        if (i == j) continue;
                        float4 aim = E[first_target + i];
                        V3 to = xyz(aim) - o;
                        const float s = dot(d, to), d2 = dot(to, to);
                        if (s < 0.0f ? s * s > 1e-8f * d2 : s * s < 0.999f * (d2 - aim.w * aim.w) - 1e-4f * d2)
                            continue;
                        float omc = sphere_omc(aim.w, d2);
                        within += (1.0f - dot(d, to / length(to)) <= omc) ? 1.0f : 0.0f;
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

            // A light path deposits and splats only past a delta bounce: one
            // whose first ray can't meet a delta surface in the FP32 walk
            // ends at its first hit with nothing to show, so it isn't traced.
            if (TREE == TREE_FP32 && alive && S.delta_count != NO_DELTA_CUT && !may_meet_delta(S, o, d))
                alive = false;
            float wavelength = 0.0f;
            uint32_t bounces = 0u;
            V3 carried = {1.0f, 1.0f, 1.0f};
            NoCount nc;
            for (uint32_t depth = 0u; alive && depth < P.max_bounces; depth++) {
                Hit hit = trace<false>(S, o, d, INF, stack, nc);
                if (hit.prim == NONE) break;
                // The hit primitive's material alone ends the path at a
                // light, or at a surface other than a delta one before any
                // delta bounce, so those paths skip building the surface.
                Mat m = load_mat(S, __float_as_uint(S.prims[(size_t)hit.prim * 4].y));
                if (m.kind == LIGHT || (bounces == 0u && !is_delta(m))) break;
                Surface s = surface<TEXTURES>(S, o, d, hit);
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
        if (wave_lane == leader) base = atomicAdd(&M.counts[PHOTON_TABLE], (uint32_t)__popcll(mask));
        base = __shfl(base, (int)leader);
        if (!pending) continue;
        uint32_t i = base + (uint32_t)__popcll(mask & ((1ull << wave_lane) - 1ull));
        if (i >= M.capacity) continue;
        V3 q = dep_p / P.photon_cell;
        uint32_t h = photon_cell_hash(to_i32(floorf(q.x)), to_i32(floorf(q.y)), to_i32(floorf(q.z)));
        uint32_t rank = atomicAdd(&M.counts[h], 1u);
        M.deposits[(size_t)i * 3] = make_float4(dep_p.x, dep_p.y, dep_p.z, __uint_as_float(h));
        M.deposits[(size_t)i * 3 + 1] = make_float4(dep_beta.x, dep_beta.y, dep_beta.z, __uint_as_float(rank));
        M.deposits[(size_t)i * 3 + 2] = make_float4(-dep_d.x, -dep_d.y, -dep_d.z, 0.0f);
    }
}

template <bool TEXTURES, bool DISPERSION, int TREE, int SW>
__global__ __launch_bounds__(BLOCK_THREADS) __attribute__((amdgpu_waves_per_eu(SW))) void k_light_trace(
    DevScene S, Params P, const float4* __restrict__ E, uint32_t* splats, PhotonMap M, uint32_t* spill,
    uint32_t slots, uint32_t photons_n) {
    __shared__ uint2 lds[MAIN_LDS_DEPTH * BLOCK_THREADS];
    auto stack = BlockStack<MAIN_LDS_DEPTH, TREE>::make(lds, spill, blockIdx.x * BLOCK_THREADS + threadIdx.x, slots);
    light_paths<TEXTURES, DISPERSION>(S, P, E, splats, M, stack, photons_n, blockIdx.x, gridDim.x);
}

// Several jobs' light tracing in one launch (rm_frame_batch): the grid is
// each job's light-tracing grid, one job after another, and each block runs
// its share of its job's light paths exactly as k_light_trace's block of the
// same index would (the same photon ids, spill lanes and deposits). The job
// table is the kernel argument, read by scalar loads. The kernel is built
// for the union of the jobs' features: a job's paths come out the same in a
// kernel with more of them.
template <bool TEXTURES, bool DISPERSION, int TREE>
__global__ __launch_bounds__(BLOCK_THREADS) __attribute__((amdgpu_waves_per_eu(5))) void k_light_trace_batch(
    LightTable T) {
    __shared__ uint2 lds[MAIN_LDS_DEPTH * BLOCK_THREADS];
    uint32_t j = 0;
    while (j + 1u < T.n && blockIdx.x >= T.begin[j + 1u]) ++j;
    const LightJob& J = T.job[j];
    const uint32_t block = blockIdx.x - T.begin[j], blocks = T.begin[j + 1u] - T.begin[j];
    auto stack = BlockStack<MAIN_LDS_DEPTH, TREE>::make(lds, J.spill, block * BLOCK_THREADS + threadIdx.x, J.slots);
    light_paths<TEXTURES, DISPERSION>(J.S, J.P, J.E, J.splats, J.M, stack, J.count, block, blocks);
}

// ------------------------------------------------------------------ photon sort
// An exclusive scan over a block's 256 threads of one value each; every
// thread gets the sum of the values before its own, and `total` the block's.
RM_D uint32_t block_exclusive_scan(uint32_t v, uint32_t* wave_totals, uint32_t& total) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t incl = v;
#pragma unroll
    for (uint32_t off = 1u; off < 64u; off <<= 1) {
        const uint32_t o = __shfl_up(incl, off);
        if (lane >= off) incl += o;
    }
    if (lane == 63u) wave_totals[wave] = incl;
    __syncthreads();
    uint32_t before = 0u;
    total = 0u;
#pragma unroll
    for (uint32_t w = 0; w < BLOCK_THREADS / 64; ++w) {
        before += w < wave ? wave_totals[w] : 0u;
        total += wave_totals[w];
    }
    return before + incl - v;
}

// Each block scans one group of PHOTON_GROUP_CELLS cell counts, 16 cells a
// thread: every cell's first slot within its group, and the group's total.
static_assert(PHOTON_GROUP_CELLS == BLOCK_THREADS * 16 && PHOTON_GROUPS == BLOCK_THREADS, "scan layout");
__global__ __launch_bounds__(BLOCK_THREADS) void k_photon_scan(const uint32_t* __restrict__ counts,
                                                               uint32_t* __restrict__ starts) {
    __shared__ uint32_t wave_totals[BLOCK_THREADS / 64];
    const uint32_t base = blockIdx.x * PHOTON_GROUP_CELLS + threadIdx.x * 16u;
    const uint4* in = reinterpret_cast<const uint4*>(counts + base);
    uint4 c[4];
    uint32_t run = 0u;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint4 v = in[k];
        c[k] = make_uint4(run, run + v.x, run + v.x + v.y, run + v.x + v.y + v.z);
        run += v.x + v.y + v.z + v.w;
    }
    uint32_t total;
    const uint32_t before = block_exclusive_scan(run, wave_totals, total);
    uint4* out = reinterpret_cast<uint4*>(starts + base);
#pragma unroll
    for (int k = 0; k < 4; ++k) out[k] = make_uint4(c[k].x + before, c[k].y + before, c[k].z + before, c[k].w + before);
    if (threadIdx.x == 0) starts[PHOTON_TABLE + blockIdx.x] = total;
}

// Every deposited photon to its cell's run, at its rank: the runs together
// hold exactly the photons in slots below min(claimed, capacity). Each block
// first scans the groups' counts in LDS for their first slots (block 0 keeps
// them for the gather's walk).
__global__ __launch_bounds__(BLOCK_THREADS) void k_photon_sort(PhotonMap M) {
    __shared__ uint32_t wave_totals[BLOCK_THREADS / 64];
    __shared__ uint32_t group_first[PHOTON_GROUPS];
    uint32_t total;
    const uint32_t first = block_exclusive_scan(M.starts[PHOTON_TABLE + threadIdx.x], wave_totals, total);
    group_first[threadIdx.x] = first;
    if (blockIdx.x == 0) M.starts[PHOTON_TABLE + PHOTON_GROUPS + threadIdx.x] = first;
    __syncthreads();
    const uint32_t claimed = M.counts[PHOTON_TABLE];
    const uint32_t n = claimed < M.capacity ? claimed : M.capacity;
    for (uint32_t i = blockIdx.x * BLOCK_THREADS + threadIdx.x; i < n; i += gridDim.x * BLOCK_THREADS) {
        const float4 a = M.deposits[(size_t)i * 3], b = M.deposits[(size_t)i * 3 + 1];
        const float4 d = M.deposits[(size_t)i * 3 + 2];
        const uint32_t h = __float_as_uint(a.w);
        const uint32_t at = group_first[h / PHOTON_GROUP_CELLS] + M.starts[h] + __float_as_uint(b.w);
        M.positions[at] = make_float4(a.x, a.y, a.z, 0.0f);
        M.powers[(size_t)at * 2] = make_float4(b.x, b.y, b.z, 0.0f);
        M.powers[(size_t)at * 2 + 1] = d;
    }
}

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

// ------------------------------------------------------------------ photon gather
// photon_gather() (src/shaders.mjs:1577-1626), once the photon map is sorted:
//   k_gather_begin  every pixel's camera ray to its first hit. A path whose
//                   first hit is a delta surface bounces there and goes on.
//                   A wave where many paths go on carries each to its
//                   gather point (its first non-delta surface), and when
//                   every lane's photons are few, gathers there, a lane a
//                   point, as the reference does; when some lane has many,
//                   its points are queued. A wave where few paths go on
//                   queues them instead, so that they don't hold the wave
//                   through their bounces and walks while its other lanes
//                   idle;
//   k_gather_paths  each queued path on to its gather point, a thread a path,
//                   the point kept in the path's slot;
//   k_gather_walk   the queued points' photons, eight lanes to a point.
// The queue has a 6-float4 slot a pixel: points from the front (count[0] of
// them) and paths from the back (count[1], slot W·H - 1 - j for the j-th).
// A path: (o, pixel), (d, rng state), (throughput, next depth). A gather
// point: (p, pixel), (n, material), (ng, 0), (base, 0), (wo, 0), (throughput,
// 0); a path that ends without one leaves GATHER_DEAD as its pixel. Every
// pixel's gathered[] is written once.
constexpr uint32_t GATHER_DEAD = 0xffffffffu;
// Lanes of a wave whose paths go on past their first hit, from which the
// wave carries them on itself rather than queueing them.
constexpr uint32_t GATHER_DENSE_LANES = 48u;
// The most photons any lane of such a wave may have to visit for the wave
// to gather a lane a point; past it, the eight-lane walk does.
constexpr uint32_t GATHER_LANE_PHOTONS = 256u;

// Where a camera path gathers, with what gather() reads of it.
struct GatherPoint { V3 p, n, ng, base, wo, throughput; uint32_t material; };

// What a gather path reaches in one bounce.
enum class GatherStep { End, Bounce, Point };

// One bounce of a gather path at depth `depth` from o along d: at a
// non-delta surface past the first hit, its gather point q; at a delta
// surface, its bounce (o, d and throughput carry on).
template <bool TEXTURES, class STACK>
RM_D GatherStep gather_bounce(const Ctx& C, Rng& g, STACK& stack, uint32_t depth, V3& o, V3& d, V3& throughput,
                              GatherPoint& q) {
    const DevScene& S = C.S;
    NoCount nc;
    Hit hit = trace<false>(S, o, d, INF, stack, nc);
    if (hit.prim == NONE) return GatherStep::End;
    // The hit primitive's material alone decides whether the path ends
    // here (a light, or any surface but a delta one at the first hit), so
    // most camera rays stop before their shading point is built.
    const uint32_t material = __float_as_uint(S.prims[(size_t)hit.prim * 4].y);
    Mat m = load_mat(S, material);
    if (m.kind == LIGHT || (depth == 0u && !is_delta(m))) return GatherStep::End;
    Surface s = surface<TEXTURES>(S, o, d, hit);
    V3 base = albedo<TEXTURES>(S, m, s.p, s.uv);
    if (!is_delta(m)) {
        q = {s.p, s.n, s.ng, base, -d, throughput, material};
        return GatherStep::Point;
    }
    BsdfSample bs = sample_bsdf(g, m, base, s, -d);
    if (!bs.ok) return GatherStep::End;
    throughput = throughput * bs.weight;
    o = offset(s, bs.wi);
    d = bs.wi;
    return GatherStep::Bounce;
}

// A gather path from depth `depth` on to its gather point q; false when it
// ends without one.
template <bool TEXTURES, class STACK>
RM_D bool gather_path(const Ctx& C, Rng& g, STACK& stack, uint32_t depth, V3 o, V3 d, V3 throughput,
                      GatherPoint& q) {
    for (; depth < C.P->max_bounces; depth++) {
        const GatherStep step = gather_bounce<TEXTURES>(C, g, stack, depth, o, d, throughput, q);
        if (step != GatherStep::Bounce) return step == GatherStep::Point;
    }
    return false;
}

// The eight cells a gather point at p reads (the reference's cell and its
// nearer neighbour along each axis, in its order): cell k's hash-table
// entry, for k = 0 .. 7.
RM_D uint32_t gather_cell(const Params& P, V3 p, uint32_t k) {
    const V3 here = p / P.photon_cell;
    const int32_t cx = to_i32(floorf(here.x)), cy = to_i32(floorf(here.y)), cz = to_i32(floorf(here.z));
    const int32_t sx = (here.x - floorf(here.x)) >= 0.5f ? 1 : -1;
    const int32_t sy = (here.y - floorf(here.y)) >= 0.5f ? 1 : -1;
    const int32_t sz = (here.z - floorf(here.z)) >= 0.5f ? 1 : -1;
    return photon_cell_hash(cx + (int32_t)(k & 1u) * sx, cy + (int32_t)((k >> 1u) & 1u) * sy,
                            cz + (int32_t)((k >> 2u) & 1u) * sz);
}
// The sorted slot of cell h's newest photon (its run ends there), given its count.
RM_D uint32_t cell_last(const PhotonMap& M, uint32_t h, uint32_t count) {
    return M.starts[PHOTON_TABLE + PHOTON_GROUPS + h / PHOTON_GROUP_CELLS] + M.starts[h] + count - 1u;
}
// One photon's term of gather()'s sum at q (zero when it is out of reach).
RM_D V3 photon_term(const PhotonMap& M, uint32_t at, const GatherPoint& q, const Mat& m, float r, float r2) {
    const float4 pos = M.positions[at];
    V3 off = xyz(pos) - q.p;
    if (dot(off, off) < r2 && fabsf(dot(off, q.ng)) < 0.25f * r) {
        const V3 wi = xyz(M.powers[(size_t)at * 2 + 1]);
        float cos_in = dot(q.n, wi);
        if (cos_in > 0.0f && dot(q.ng, wi) > 0.0f) {
            Eval ev = eval_bsdf(m, q.base, q.n, q.wo, wi);
            return xyz(M.powers[(size_t)at * 2]) * (ev.f / cos_in);
        }
    }
    return {0.0f, 0.0f, 0.0f};
}

// The eight cells of a gather point: each run's newest sorted slot and
// length, and how many photons the reference visits in all (at most 4096).
struct PointCells { uint32_t last[8], count[8], total; };
RM_D PointCells point_cells(const Params& P, const PhotonMap& M, V3 p) {
    PointCells c;
    uint32_t all = 0u;
#pragma unroll
    for (uint32_t k = 0u; k < 8u; k++) {
        const uint32_t h = gather_cell(P, p, k);
        c.count[k] = M.counts[h];
        c.last[k] = cell_last(M, h, c.count[k]);
        all += c.count[k];
    }
    c.total = all < 4096u ? all : 4096u;
    return c;
}

// gather() (src/shaders.mjs:1534-1570) by one lane, in the reference's order:
// the eight cells in turn, each newest deposit first (the reference's lists
// are pushed at the head), at most 4096 photons in all. The run's slots are
// known up front, so unrolled iterations load their photons together. The
// cells move down one place a step, so the current one is always at index 0:
// indexed only by constants, the arrays stay in registers.
RM_D V3 gather_lane(const Params& P, const PhotonMap& M, const GatherPoint& q, PointCells c, const Mat& m) {
    const float r = P.photon_radius, r2 = r * r;
    V3 sum = {0, 0, 0};
    uint32_t visited = 0u;
    for (uint32_t k = 0u; k < 8u && visited < 4096u; k++) {
        const uint32_t take = c.count[0] < 4096u - visited ? c.count[0] : 4096u - visited;
        const uint32_t last = c.last[0];
#pragma unroll 4
        for (uint32_t i = 0u; i < take; i++) sum = sum + photon_term(M, last - i, q, m, r, r2);
        visited += take;
#pragma unroll
        for (uint32_t j = 0u; j < 7u; j++) {
            c.count[j] = c.count[j + 1u];
            c.last[j] = c.last[j + 1u];
        }
    }
    return sum / (PI * r2);
}

// The gathered light a pixel's point adds (w = 1), or zeros.
RM_D float4 gathered_light(V3 throughput, V3 radiance) {
    const V3 light = throughput * radiance;
    return finite3(light) ? make_float4(light.x, light.y, light.z, 1.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Every pixel's camera ray (the reference's gather stream) to its first hit,
// and on (see above).
template <bool TEXTURES, int TREE, int SW>
__global__ __launch_bounds__(BLOCK_THREADS) __attribute__((amdgpu_waves_per_eu(SW))) void k_gather_begin(DevScene S, Params P, PhotonMap M, GatherQueue Q,
                                                                float4* __restrict__ gathered, uint32_t* spill,
                                                                uint32_t slots) {
    __shared__ uint2 lds[GATHER_LDS_DEPTH * BLOCK_THREADS];
    uint32_t x, y;
    tile_pixel(x, y);
    const bool inside = x < P.width && y < P.height;
    const Ctx C{S, &P};
    const uint32_t index = inside ? y * P.width + x : 0u;
    auto stack = BlockStack<GATHER_LDS_DEPTH, TREE>::make(lds, spill, index, slots);
    bool go = false;
    Rng g;
    V3 o = {0, 0, 0}, d = {0, 0, 0}, throughput = {1, 1, 1};
    GatherPoint q = {};
    if (inside && P.max_bounces > 0u) {
        g.qmc = false;
        g.state = pcg(index ^ pcg(P.sample_offset * 2654435769u + 4099u));
        float jx = g.rand();
        float jy = g.rand();
        Ray ray = camera_ray(P, g, {(float)x + jx, (float)y + jy}, true);
        o = ray.o;
        d = ray.d;
        // A primary ray whose FP32 walk can't meet a delta surface ends at
        // its first hit, whatever that is: no need to find it.
        if (TREE != TREE_FP32 || S.delta_count == NO_DELTA_CUT || may_meet_delta(S, o, d))
            go = gather_bounce<TEXTURES>(C, g, stack, 0u, o, d, throughput, q) == GatherStep::Bounce;
    }
    if ((uint32_t)__popcll(__ballot(go)) >= GATHER_DENSE_LANES) {
        const bool found = go && gather_path<TEXTURES>(C, g, stack, 1u, o, d, throughput, q);
        PointCells cells;
        cells.total = 0u;
        if (found) cells = point_cells(P, M, q.p);
        // A lane a point while every lane's walk is short; else to the queue.
        if (__ballot(cells.total > GATHER_LANE_PHOTONS) == 0ull) {
            if (found) gathered[index] = gathered_light(q.throughput, gather_lane(P, M, q, cells, load_mat(S, q.material)));
            else if (inside) gathered[index] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            return;
        }
        const uint32_t at = wave_slot(&Q.count[0], found);
        if (found) {
            float4* slot = Q.slots + (size_t)at * GATHER_SLOT_VEC4S;
            slot[0] = make_float4(q.p.x, q.p.y, q.p.z, __uint_as_float(index));
            slot[1] = make_float4(q.n.x, q.n.y, q.n.z, __uint_as_float(q.material));
            slot[2] = make_float4(q.ng.x, q.ng.y, q.ng.z, 0.0f);
            slot[3] = make_float4(q.base.x, q.base.y, q.base.z, 0.0f);
            slot[4] = make_float4(q.wo.x, q.wo.y, q.wo.z, 0.0f);
            slot[5] = make_float4(q.throughput.x, q.throughput.y, q.throughput.z, 0.0f);
        } else if (inside) {
            gathered[index] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        return;
    }
    const uint32_t at = wave_slot(&Q.count[1], go);
    if (go) {
        float4* slot = Q.slots + (size_t)(Q.px - 1u - at) * GATHER_SLOT_VEC4S;
        slot[0] = make_float4(o.x, o.y, o.z, __uint_as_float(index));
        slot[1] = make_float4(d.x, d.y, d.z, __uint_as_float(g.state));
        slot[2] = make_float4(throughput.x, throughput.y, throughput.z, __uint_as_float(1u));
    } else if (inside) {
        gathered[index] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

// The queued paths on to their gather points, a thread a path, each point
// written over its path's slot.
template <bool TEXTURES, int TREE, int SW>
__global__ __launch_bounds__(BLOCK_THREADS) __attribute__((amdgpu_waves_per_eu(SW))) void k_gather_paths(DevScene S, Params P, GatherQueue Q,
                                                                float4* __restrict__ gathered, uint32_t* spill,
                                                                uint32_t slots) {
    __shared__ uint2 lds[GATHER_LDS_DEPTH * BLOCK_THREADS];
    const Ctx C{S, &P};
    const uint32_t queued = Q.count[1];
    const uint32_t lane = blockIdx.x * BLOCK_THREADS + threadIdx.x;
    auto stack = BlockStack<GATHER_LDS_DEPTH, TREE>::make(lds, spill, lane, slots);
    for (uint32_t j = lane; j < queued; j += gridDim.x * BLOCK_THREADS) {
        float4* slot = Q.slots + (size_t)(Q.px - 1u - j) * GATHER_SLOT_VEC4S;
        const float4 a = slot[0], b = slot[1], c = slot[2];
        const uint32_t pixel = __float_as_uint(a.w);
        Rng g;
        g.qmc = false;
        g.state = __float_as_uint(b.w);
        GatherPoint q;
        if (gather_path<TEXTURES>(C, g, stack, __float_as_uint(c.w), xyz(a), xyz(b), xyz(c), q)) {
            slot[0] = make_float4(q.p.x, q.p.y, q.p.z, __uint_as_float(pixel));
            slot[1] = make_float4(q.n.x, q.n.y, q.n.z, __uint_as_float(q.material));
            slot[2] = make_float4(q.ng.x, q.ng.y, q.ng.z, 0.0f);
            slot[3] = make_float4(q.base.x, q.base.y, q.base.z, 0.0f);
            slot[4] = make_float4(q.wo.x, q.wo.y, q.wo.z, 0.0f);
            slot[5] = make_float4(q.throughput.x, q.throughput.y, q.throughput.z, 0.0f);
        } else {
            slot[0] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(GATHER_DEAD));
            gathered[pixel] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    }
}

// gather() for the queued points, eight lanes to a point: the point's eight
// cells, one a lane, are runs of the sorted photons. The first 4096 photons
// of the runs in the reference's order are laid end to end and the eight
// lanes take every eighth, so they load side by side rather than one link at
// a time. The lanes' partial sums are added in a fixed tree; the photons'
// order within a cell follows their deposits, which the reference leaves to
// the order its atomics land in.
template <int SW>
__global__ __launch_bounds__(BLOCK_THREADS) __attribute__((amdgpu_waves_per_eu(SW))) void k_gather_walk(DevScene S, Params P, PhotonMap M, GatherQueue Q,
                                                               float4* __restrict__ gathered) {
    const uint32_t front = Q.count[0], points = front + Q.count[1];
    const uint32_t sub = __lane_id() & 7u;
    const uint32_t stride = gridDim.x * BLOCK_THREADS / 8u;
    const float r = P.photon_radius, r2 = r * r;
    for (uint32_t v = (blockIdx.x * BLOCK_THREADS + threadIdx.x) / 8u; __ballot(v < points); v += stride) {
        bool live = false;
        GatherPoint q = {};
        uint32_t pixel = GATHER_DEAD;
        if (v < points) {
            const float4* slot = Q.slots + (size_t)(v < front ? v : Q.px - 1u - (v - front)) * GATHER_SLOT_VEC4S;
            const float4 a = slot[0];
            pixel = __float_as_uint(a.w);
            live = pixel != GATHER_DEAD;
            if (live) {
                const float4 b = slot[1], c = slot[2], e = slot[3], f = slot[4], t = slot[5];
                q = {xyz(a), xyz(b), xyz(c), xyz(e), xyz(f), xyz(t), __float_as_uint(b.w)};
            }
        }
        // This lane's cell of the point, and where its run starts in the runs
        // laid end to end (an inclusive scan over the group's eight lanes).
        uint32_t count = 0u, last = 0u;
        if (live) {
            const uint32_t h = gather_cell(P, q.p, sub);
            count = M.counts[h];
            last = cell_last(M, h, count);
        }
        uint32_t end = count;
#pragma unroll
        for (uint32_t off = 1u; off < 8u; off <<= 1) {
            const uint32_t below = __shfl_up(end, off, 8);
            if (sub >= off) end += below;
        }
        uint32_t begins[8], lasts[8];
#pragma unroll
        for (uint32_t k = 0u; k < 8u; ++k) {
            begins[k] = __shfl(end - count, (int)k, 8);
            lasts[k] = __shfl(last, (int)k, 8);
        }
        // The runs laid end to end, cut at the first 4096 photons.
        const uint32_t all = __shfl(end, 7, 8);
        const uint32_t total = all < 4096u ? all : 4096u;
        V3 sum = {0, 0, 0};
        if (live) {
            const Mat m = load_mat(S, q.material);
            for (uint32_t j = sub; j < total; j += 8u) {
                uint32_t k = 0u;
#pragma unroll
                for (uint32_t c = 1u; c < 8u; c++) k += j >= begins[c] ? 1u : 0u;
                sum = sum + photon_term(M, lasts[k] - (j - begins[k]), q, m, r, r2);
            }
        }
#pragma unroll
        for (int off = 4; off > 0; off >>= 1) {
            sum.x += __shfl_xor(sum.x, off, 8);
            sum.y += __shfl_xor(sum.y, off, 8);
            sum.z += __shfl_xor(sum.z, off, 8);
        }
        if (live && sub == 0u) gathered[pixel] = gathered_light(q.throughput, sum / (PI * r2));
    }
}

// ------------------------------------------------------------------ launchers
template <bool T, bool D, int SW>
static void launch_light_trace_v(const DevScene& S, const Params& P, const float4* emitters, uint32_t* splats,
                                 const PhotonMap& M, uint32_t* spill, uint32_t slots, dim3 grid, uint32_t count,
                                 int tree, hipStream_t stream) {
    with_tree(tree, [&](auto t) {
        hipLaunchKernelGGL((k_light_trace<T, D, decltype(t)::value, SW>), grid, dim3(BLOCK_THREADS), 0, stream, S, P,
                           emitters, splats, M, spill, slots, count);
    });
}
template <int SW>
static void launch_light_trace_w(const DevScene& S, const Params& P, const float4* emitters, uint32_t* splats,
                                 const PhotonMap& M, uint32_t* spill, uint32_t slots, dim3 grid, uint32_t count,
                                 const Features& f, hipStream_t stream) {
    if (f.textures)
        launch_light_trace_v<true, true, SW>(S, P, emitters, splats, M, spill, slots, grid, count, f.tree, stream);
    else if (f.dispersion)
        launch_light_trace_v<false, true, SW>(S, P, emitters, splats, M, spill, slots, grid, count, f.tree, stream);
    else
        launch_light_trace_v<false, false, SW>(S, P, emitters, splats, M, spill, slots, grid, count, f.tree, stream);
}

void launch_light_trace(const DevScene& S, const Params& P, const float4* emitters, uint32_t* splats,
                        const PhotonMap& M, uint32_t* spill, uint32_t slots, uint32_t count, const Features& f,
                        int side_waves, hipStream_t stream) {
    const dim3 grid(light_blocks(count));
    if (grid.x == 0) return;
    if (side_waves == 7) launch_light_trace_w<7>(S, P, emitters, splats, M, spill, slots, grid, count, f, stream);
    else launch_light_trace_w<5>(S, P, emitters, splats, M, spill, slots, grid, count, f, stream);
}

template <bool T, bool D>
static void launch_light_trace_batch_v(const LightTable& L, int tree, hipStream_t stream) {
    with_tree(tree, [&](auto t) {
        hipLaunchKernelGGL((k_light_trace_batch<T, D, decltype(t)::value>), dim3(L.begin[L.n]), dim3(BLOCK_THREADS), 0,
                           stream, L);
    });
}

void launch_light_trace_batch(const LightTable& L, const Features& f, hipStream_t stream) {
    if (L.begin[L.n] == 0) return;
    if (f.textures) launch_light_trace_batch_v<true, true>(L, f.tree, stream);
    else if (f.dispersion) launch_light_trace_batch_v<false, true>(L, f.tree, stream);
    else launch_light_trace_batch_v<false, false>(L, f.tree, stream);
}

void launch_photon_sort(const PhotonMap& M, hipStream_t stream) {
    hipLaunchKernelGGL(k_photon_scan, dim3(PHOTON_GROUPS), dim3(BLOCK_THREADS), 0, stream, M.counts, M.starts);
    // The deposit count is known on the device only: a grid for the
    // capacity, at most 1024 blocks, each thread taking every grid's-worth.
    uint32_t blocks = (M.capacity + BLOCK_THREADS - 1) / BLOCK_THREADS;
    blocks = blocks < 1u ? 1u : blocks < 1024u ? blocks : 1024u;
    hipLaunchKernelGGL(k_photon_sort, dim3(blocks), dim3(BLOCK_THREADS), 0, stream, M);
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

template <int SW>
static void launch_photon_gather_w(const DevScene& S, const Params& P, const PhotonMap& M, const GatherQueue& Q,
                                   float4* gathered, uint32_t* spill, uint32_t slots, const Features& f,
                                   hipStream_t stream) {
    const uint32_t px = P.width * P.height;
    const dim3 tiles((P.width + 15) / 16, (P.height + 15) / 16);
    // The queue's length is known on the device only: grids for at most
    // every pixel, each thread (or group of eight) taking every grid's-worth.
    uint32_t path_blocks = (px + BLOCK_THREADS - 1) / BLOCK_THREADS;
    path_blocks = path_blocks < 2048u ? path_blocks : 2048u;
    uint32_t walk_blocks = (px + BLOCK_THREADS / 8 - 1) / (BLOCK_THREADS / 8);
    walk_blocks = walk_blocks < 2048u ? walk_blocks : 2048u;
    launch_fill_u32(Q.count, 0u, 2, stream);
    auto go = [&](auto begin, auto paths) {
        hipLaunchKernelGGL(begin, tiles, dim3(BLOCK_THREADS), 0, stream, S, P, M, Q, gathered, spill, slots);
        hipLaunchKernelGGL(paths, dim3(path_blocks), dim3(BLOCK_THREADS), 0, stream, S, P, Q, gathered, spill,
                           slots);
    };
    with_tree(f.tree, [&](auto t) {
        constexpr int TREE = decltype(t)::value;
        if (f.textures) go(k_gather_begin<true, TREE, SW>, k_gather_paths<true, TREE, SW>);
        else go(k_gather_begin<false, TREE, SW>, k_gather_paths<false, TREE, SW>);
    });
    hipLaunchKernelGGL(k_gather_walk<SW>, dim3(walk_blocks), dim3(BLOCK_THREADS), 0, stream, S, P, M, Q, gathered);
}
void launch_photon_gather(const DevScene& S, const Params& P, const PhotonMap& M, const GatherQueue& Q,
                          float4* gathered, uint32_t* spill, uint32_t slots, const Features& f, int side_waves,
                          hipStream_t stream) {
    if (side_waves == 7) launch_photon_gather_w<7>(S, P, M, Q, gathered, spill, slots, f, stream);
    else launch_photon_gather_w<5>(S, P, M, Q, gathered, spill, slots, f, stream);
}

}  // namespace rm
