// Launch interface between the C-ABI layer (api.hip) and the kernels
// (kernels.hip), with the block shape and traversal-stack depths.
#pragma once
#include "shading.hpp"

namespace rm {

// 256 threads: four wave64s covering a 16×16 pixel tile.
constexpr int BLOCK_THREADS = 256;
// Traversal stack entries kept in LDS per thread (the rest of the
// reference's 64 spill to global memory), 8 bytes each: 16 KiB of LDS per
// block of the path pass. Sixteen entries (32 KiB) measured the same on a
// 200-sphere scene and 1.4% slower on a 2M-triangle one: the LDS that eight
// leave free lets the side stream's light tracing and gather share CUs with
// the path pass.
constexpr int MAIN_LDS_DEPTH = 8;
constexpr int TRACE_LDS_DEPTH = 8;
constexpr int GATHER_LDS_DEPTH = 8;
// Spill entries (8 bytes each) per lane: the rest of the 64 beyond the
// shallowest LDS stack (the compacted path kernel keeps 4 entries in LDS).
constexpr uint32_t SPILL_ENTRIES = STACK_LIMIT - 4u;

// What a frame's scene and options need; each combination is its own kernel.
// tree: the tree the traversals walk (TREE_FP32, TREE_QUANTIZED, TREE_WIDE) in
// the per-pixel path pass, raycast, light tracing and gather; the compact and
// persistent path passes walk the FP32 tree.
struct Features { bool volumes, textures, environment, dispersion; int tree; };

struct LaunchMain {
    DevScene S;
    Params P;
    float4* accum;
    float4* galb;
    float4* gnorm;
    float4* comp;                // non-null: compensated sums, 3 float4 of rounding error a pixel
    uint32_t* spill;
    uint32_t slots;
    uint32_t persistent_blocks;  // > 0: k_main_persistent with this grid
    uint32_t* counter;           // k_main_persistent's tile counter, zeroed before the launch
    float4* staging;             // non-null: k_main_compact, 3 float4 per thread of its grid
    int tree;                    // the tree the per-pixel pass walks
    int waves;                   // waves per SIMD the one-thread-per-pixel pass is compiled for: 4 or 5
    bool sorted = false;         // with staging: k_main_compact sorting live paths by key between bounces
    int stack_depth = 8;         // k_main at four waves: traversal stack entries in LDS, 8 or 16
    int block_shape = 0;         // k_main at four waves: BLOCK_SQUARES (16×16) or BLOCK_ROWS (64×4)
    uint32_t* split = nullptr;   // non-null: the split schedule's queue of W·H slots, then its two counters
    uint32_t split_depth = 2;    // bounces the split schedule's head kernel runs before queueing a path
    uint32_t split_blocks = 0;   // the tail kernel's persistent grid
};
// Workgroups go to the XCDS XCDs round-robin by linear block id. A stencil
// pass's grid (stencil_pixel) has its 16×16 tiles rounded up to a multiple
// of XCDS blocks.
constexpr uint32_t XCDS = 8u;
__host__ __device__ inline uint32_t stencil_blocks(uint32_t width, uint32_t height) {
    uint32_t tiles = ((width + 15u) / 16u) * ((height + 15u) / 16u);
    return (tiles + XCDS - 1u) / XCDS * XCDS;
}
// The split schedule's buffer for px pixels: the queue (SLOT_WORDS words a
// path, the path slots' layout), then two counters at this word offset.
__host__ __device__ inline size_t split_count_offset(size_t px) { return (px * 27 + 63) / 64 * 64; }
inline size_t split_bytes(size_t px) { return (split_count_offset(px) + 64) * 4; }
// The tail's persistent grid on a device of `cus` compute units: as many
// 256-thread blocks a CU as the kernels' waves per SIMD (four SIMDs a CU).
inline uint32_t split_tail_blocks(int cus, int waves) { return (uint32_t)(cus > 0 ? cus : 1) * (uint32_t)waves; }

void launch_main(const LaunchMain& L, const Features& f, hipStream_t stream);
// The one-thread-per-pixel pass compiled for five waves per SIMD (main5.hip).
void launch_main_waves5(const LaunchMain& L, const Features& f, hipStream_t stream);
// k_main at four waves with 16 stack entries in LDS, or with 64×4 row blocks (main_shapes.hip).
void launch_main_stack16(const LaunchMain& L, const Features& f, hipStream_t stream);
void launch_main_rows(const LaunchMain& L, const Features& f, hipStream_t stream);
// The persistent and compacted schedules' launches (launch_main picks them).
void launch_main_persistent(const LaunchMain& L, const Features& f, hipStream_t stream);
void launch_main_compact(const LaunchMain& L, const Features& f, hipStream_t stream);
// The split schedule: a head kernel over pixels, then a persistent tail over
// the paths still alive after split_depth bounces (split.hip).
void launch_main_split(const LaunchMain& L, const Features& f, hipStream_t stream);
void launch_main_split5(const LaunchMain& L, const Features& f, hipStream_t stream);

// Several jobs' path passes in one launch of the per-pixel kernel, the job
// table passed by value as the kernel's argument.
constexpr uint32_t MAX_BATCH = 8;
struct BatchJob {
    DevScene S;
    Params P;
    float4* accum;
    float4* galb;
    float4* gnorm;
    float4* comp;        // non-null: compensated sums
    uint32_t* spill;
    uint32_t slots;
    uint32_t tiles_x;    // 16×16 tiles across the job's image
};
struct BatchTable {
    BatchJob job[MAX_BATCH];
    uint32_t begin[MAX_BATCH + 1];   // job j's tiles are [begin[j], begin[j + 1]) of the grid
    uint32_t n;
};
static_assert(sizeof(BatchTable) <= 4096, "the job table travels as the kernel argument");
// f: the union of the jobs' features (a job renders the same in a kernel with more of them).
void launch_main_batch(const BatchTable& B, const Features& f, hipStream_t stream);

// The wavefront path pass's buffers (RM_SCHEDULE_WAVEFRONT), for n paths (the
// frame's pixels): path state, closest hits, pending shadow rays, each pixel's
// sums, the live queues of this bounce and the next, the shadow and done
// queues, and their four counters.
constexpr int WF_STATE_WORDS = 27;
constexpr int WF_NEE_WORDS = 15;
constexpr int WF_DONE_WORDS = 11;
struct Wavefront {
    uint32_t* state[2];   // [WF_STATE_WORDS][n] each: the live paths of this bounce and the next, in queue order
    uint4* hits;          // [n] the live paths' closest hits, in queue order
    uint32_t* nee;        // [WF_NEE_WORDS][n] shadow rays and where their path's result went
    uint32_t* done;       // [WF_DONE_WORDS][n] the paths that ended this bounce
    float4* sums;         // [3][n] each pixel's sums
    uint32_t* counts;     // [4]: live now, live next, shadow rays, done
    uint32_t n;
};
// Bytes of device memory a Wavefront of n paths needs, each of its seven
// buffers starting on a 256-byte boundary.
inline size_t wavefront_bytes(size_t n) {
    return n * ((2 * WF_STATE_WORDS + WF_NEE_WORDS + WF_DONE_WORDS) * 4 + 16 + 48) + 7 * 256;
}

struct LaunchWavefront {
    DevScene S;
    Params P;
    Wavefront W;
    float4* accum;
    float4* galb;
    float4* gnorm;
    float4* comp;       // non-null: compensated sums
    uint32_t* spill;    // traversal spill for blocks × BLOCK_THREADS slots
    uint32_t slots;
    uint32_t blocks;    // the grid of every stage: a thread per path
    int tree;
};
void launch_wavefront(const LaunchWavefront& L, const Features& f, hipStream_t stream);

// Several jobs' wavefront path passes, each stage one launch for all of them
// (rm_frame_batch with every context on RM_SCHEDULE_WAVEFRONT). The table is
// in device memory; job j's queue entries are its 256-entry chunks
// [begin[j], begin[j + 1]).
struct WfJob {
    DevScene S;
    Params P;
    Wavefront W;
    float4* accum;
    float4* galb;
    float4* gnorm;
    float4* comp;       // non-null: compensated sums
    uint32_t* spill;    // traversal spill for `slots` entries (at least W.n)
    uint32_t slots;
};
struct WfTable {
    WfJob job[MAX_BATCH];
    uint32_t begin[MAX_BATCH + 1];
    uint32_t n;
};
// The launch sequence's shape: samples per frame and bounce stages (the
// most any job has), the persistent grid, the tree. Captured in a hipGraph,
// one sequence serves every batch of this shape.
struct WfBatchShape {
    uint32_t samples;
    uint32_t bounces;   // extend/shade/shadow/finish rounds: max_bounces + 1 of the deepest job
    uint32_t blocks;
    int tree;
};
// f: the union of the jobs' features.
void launch_wavefront_batch(const WfTable* table, const WfBatchShape& B, const Features& f, hipStream_t stream);
// How many 256-thread blocks of the persistent path kernel fit on one CU.
int main_resident_blocks_per_cu(const Features& f);
void launch_raycast(const DevScene& S, const float4* rays, uint32_t n, uint4* hits, bool any_hit, int tree,
                    uint32_t* spill, uint32_t slots, uint2* counters, hipStream_t stream);
// The reference's light paths (and photon slots) a frame: one a sample, at
// most PHOTON_CAPACITY (src/shaders.mjs:1495-1497).
__host__ __device__ inline uint32_t photon_count(const Params& P) {
    const uint64_t n = (uint64_t)P.width * P.height * P.samples_per_frame;
    return n < PHOTON_CAPACITY ? (uint32_t)n : PHOTON_CAPACITY;
}

// A frame's photon map. Light tracing deposits each photon in the slot its
// wave claims and counts it in its hash cell (the reference's 2^20-cell
// grid, src/shaders.mjs:1499-1528); launch_photon_sort then lays the photons
// out cell by cell, so a gather reads each cell's photons as one contiguous
// run rather than following a linked list one dependent load at a time.
constexpr uint32_t PHOTON_GROUP_CELLS = 4096;                          // cells one scan block covers
constexpr uint32_t PHOTON_GROUPS = PHOTON_TABLE / PHOTON_GROUP_CELLS;  // 256
struct PhotonMap {
    uint32_t* counts;    // PHOTON_TABLE + 1 words: photons in each cell; [PHOTON_TABLE] counts the slots claimed
    uint32_t* starts;    // PHOTON_TABLE + 2 · PHOTON_GROUPS words: each cell's first sorted slot within its group
                         // of PHOTON_GROUP_CELLS cells, each group's photon count, then each group's first slot
    float4* deposits;    // capacity × 3, slot order: (position, cell), (power, rank within the cell), (to the light, 0)
    float4* positions;   // capacity, cell order: (position, 0)
    float4* powers;      // capacity × 2, cell order: (power, 0), (to the light, 0)
    uint32_t capacity;   // photon slots; deposits past it are dropped, as the reference drops them
};
// Bytes of the photon map's buffers other than the counts: starts, and the
// deposits and their sorted copy for `capacity` photons.
inline uint64_t photon_map_bytes(uint64_t capacity) {
    return (uint64_t)(PHOTON_TABLE + 2 * PHOTON_GROUPS) * 4 + capacity * 96;
}

// Light tracing: count light paths, photon ids 0 .. count - 1, each depositing
// at most one photon into M (whose counts start at zero); the grid is at most
// LIGHT_LANES threads, and `spill` has room for that many lanes.
constexpr uint32_t LIGHT_LANES = 1u << 20;
// side_waves (5 or 7): the waves per SIMD light tracing and the gather are
// compiled for beside the path pass (side_waves()).
void launch_light_trace(const DevScene& S, const Params& P, const float4* emitters, uint32_t* splats,
                        const PhotonMap& M, uint32_t* spill, uint32_t slots, uint32_t count, const Features& f,
                        int side_waves, hipStream_t stream);
// Light tracing's and the gather's waves per SIMD beside a path pass of
// `path_waves`: more than it, so that a wave of theirs fits where a path-pass
// wave retires — five (96 VGPRs) beside four, seven (72 VGPRs) beside five.
inline int side_waves(int path_waves) { return path_waves == 5 ? 7 : 5; }
// Blocks of light tracing's grid for `count` light paths: one path a thread
// up to LIGHT_LANES, then each thread takes several.
inline uint32_t light_blocks(uint32_t count) {
    const uint32_t lanes = count < LIGHT_LANES ? count : LIGHT_LANES;
    return (lanes + BLOCK_THREADS - 1) / BLOCK_THREADS;
}
// Several jobs' light tracing in one launch (rm_frame_batch): job j's blocks
// are [begin[j], begin[j + 1]) of the grid, light_blocks(count) of them.
struct LightJob {
    DevScene S;
    Params P;
    const float4* E;      // the emitters
    uint32_t* splats;
    PhotonMap M;          // its counts cleared
    uint32_t* spill;
    uint32_t slots;
    uint32_t count;       // light paths
};
struct LightTable {
    LightJob job[MAX_BATCH];
    uint32_t begin[MAX_BATCH + 1];
    uint32_t n;
};
static_assert(sizeof(LightTable) <= 4096, "the job table travels as the kernel argument");
// f: the union of the jobs' features (the tree the same for all).
void launch_light_trace_batch(const LightTable& L, const Features& f, hipStream_t stream);
// Counting sort of the deposited photons by cell: an exclusive scan of the
// cell counts, then each photon copied to its cell's run at its rank.
void launch_photon_sort(const PhotonMap& M, hipStream_t stream);
void launch_resolve(const Params& P, uint4* splats, const float4* gathered, float4* accum, float4* comp,
                    hipStream_t stream);
// sum + error of the compensated sums, into out: planes accum, galb, gnorm of W·H float4.
void launch_fold(const Params& P, const float4* accum, const float4* galb, const float4* gnorm, const float4* comp,
                 float4* out, hipStream_t stream);
// The photon gather's queue of gather points (from the front) and camera
// paths (from the back, each replaced in its slot by its gather point): a
// slot of GATHER_SLOT_VEC4S float4 for each of px pixels; count[0] points
// and count[1] paths queued.
constexpr uint32_t GATHER_SLOT_VEC4S = 6;
struct GatherQueue {
    float4* slots;
    uint32_t* count;
    uint32_t px;
};
// Every pixel's gathered light into `gathered` (w = 1 where it adds any),
// once the photon map is sorted; spill has room for W·H lanes of `slots`.
void launch_photon_gather(const DevScene& S, const Params& P, const PhotonMap& M, const GatherQueue& Q,
                          float4* gathered, uint32_t* spill, uint32_t slots, const Features& f, int side_waves,
                          hipStream_t stream);
// fp16: the guide buffer holds FP16 (8 bytes a pixel) rather than FP32.
void launch_guides(const Params& P, const float4* gnorm, void* guides, bool fp16, hipStream_t stream);
// Resolve (when splats is non-null), guides and prepare in one pass, for a
// frame with 4 or more samples accumulated and plain (not compensated) sums.
void launch_finish(const Params& P, uint4* splats, const float4* gathered, float4* accum, const float4* galb,
                   const float4* gnorm, void* guides, bool fp16, float4* out, float4* albedo_out,
                   hipStream_t stream);
void launch_prepare(const Params& P, const float4* accum, const float4* galb, const void* guides, bool fp16,
                    float4* out, float4* albedo_out, hipStream_t stream);
void launch_atrous(const Params& P, int step, const void* guides, bool fp16, const float4* in, float4* out,
                   hipStream_t stream);
void launch_compose(const Params& P, const float4* albedo, const float4* in, float4* out, hipStream_t stream);
// Several jobs' denoiser passes in one launch each (rm_frame_batch): the five
// à-trous passes and compose after each job's prepare. Job j's blocks are
// [begin[j], begin[j + 1]) of each grid, stencil_blocks(width, height) of
// them. The passes go a → b → a → b → a → b, so compose reads b.
struct DenoiseJob {
    Params P;
    const void* guides;    // FP32 or FP16 guides, the same for every job of a table
    float4* a;             // prepare's output
    float4* b;
    const float4* albedo;  // prepare's albedo
    float4* out;           // the denoised image
};
struct DenoiseTable {
    DenoiseJob job[MAX_BATCH];
    uint32_t begin[MAX_BATCH + 1];
    uint32_t n;
};
static_assert(sizeof(DenoiseTable) <= 4096, "the job table travels as the kernel argument");
void launch_denoise_batch(const DenoiseTable& D, bool fp16, hipStream_t stream);
void launch_display(const Params& P, const float4* accum, const float4* denoised, uchar4* out, hipStream_t stream);
void launch_fill_u32(uint32_t* p, uint32_t value, size_t n, hipStream_t stream);

}  // namespace rm
