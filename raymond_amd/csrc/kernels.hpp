// Launch interface between the C-ABI layer (api.hip) and the kernels
// (kernels.hip), with the block shape and traversal-stack depths.
#pragma once
#include "shading.hpp"

namespace rm {

// 256 threads: four wave64s covering a 16×16 pixel tile.
constexpr int BLOCK_THREADS = 256;
// Traversal stack entries kept in LDS per thread (the rest of the
// reference's 64 spill to global memory), 8 bytes each: 32 KiB of LDS per
// block at 16, which still lets four blocks of the path pass share a CU; the
// gather and raycast kernels hold more waves, so they keep fewer.
constexpr int MAIN_LDS_DEPTH = 16;
constexpr int TRACE_LDS_DEPTH = 8;
constexpr int GATHER_LDS_DEPTH = 8;
// Spill entries (8 bytes each) per lane: the rest of the 64 beyond the
// shallowest LDS stack (the compacted path kernel keeps 4 entries in LDS).
constexpr uint32_t SPILL_ENTRIES = STACK_LIMIT - 4u;

// What a frame's scene and options need; each combination is its own kernel.
// quantized: traverse the 8-bit child-quantized tree (the per-pixel path
// pass, raycast, light tracing and gather; the compact and persistent path
// passes walk the FP32 tree).
struct Features { bool volumes, textures, environment, dispersion, quantized; };

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
    bool quantized;              // the per-pixel pass walks the quantized tree
};

void launch_main(const LaunchMain& L, const Features& f, hipStream_t stream);
// How many 256-thread blocks of the persistent path kernel fit on one CU.
int main_resident_blocks_per_cu(const Features& f);
void launch_raycast(const DevScene& S, const float4* rays, uint32_t n, uint4* hits, bool any_hit, bool quantized,
                    uint32_t* spill, uint32_t slots, uint2* counters, hipStream_t stream);
void launch_light_trace(const DevScene& S, const Params& P, const float4* emitters, uint32_t* splats, uint32_t* heads,
                        float4* photons, uint32_t* spill, uint32_t slots, uint32_t count, const Features& f,
                        hipStream_t stream);
void launch_resolve(const Params& P, uint4* splats, const float4* gathered, float4* accum, float4* comp,
                    hipStream_t stream);
// sum + error of the compensated sums, into out: planes accum, galb, gnorm of W·H float4.
void launch_fold(const Params& P, const float4* accum, const float4* galb, const float4* gnorm, const float4* comp,
                 float4* out, hipStream_t stream);
void launch_photon_gather(const DevScene& S, const Params& P, const uint32_t* heads, const float4* photons,
                          float4* gathered, uint32_t* spill, uint32_t slots, const Features& f, hipStream_t stream);
// fp16: the guide buffer holds FP16 (8 bytes a pixel) rather than FP32.
void launch_guides(const Params& P, const float4* gnorm, void* guides, bool fp16, hipStream_t stream);
void launch_prepare(const Params& P, const float4* accum, const float4* galb, const void* guides, bool fp16,
                    float4* out, float4* albedo_out, hipStream_t stream);
void launch_atrous(const Params& P, int step, const void* guides, bool fp16, const float4* in, float4* out,
                   hipStream_t stream);
void launch_compose(const Params& P, const float4* albedo, const float4* in, float4* out, hipStream_t stream);
void launch_display(const Params& P, const float4* accum, const float4* denoised, uchar4* out, hipStream_t stream);
void launch_fill_u32(uint32_t* p, uint32_t value, size_t n, hipStream_t stream);

}  // namespace rm
