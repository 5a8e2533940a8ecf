/*
 * libraymond_hip — the C ABI of Raymond's rendering path on one MI355X.
 *
 * This replaces the reference's WebGPU device boundary: the buffer uploads
 * and compute dispatches that src/Renderer.mjs issues, and the readback in
 * readPixels(). Plain pointers and sizes only; every host buffer is copied
 * during the call and never retained. Calls return RM_OK (0) or a negative
 * RM_ERR_* code; rm_last_error() then says why, in a sentence.
 *
 * Packed formats are the reference's (src/Scene.mjs:15-60, src/BVH.mjs:6-15):
 * 64-byte primitives, 32-byte BVH nodes, 96-byte materials, attribute
 * vec4s, the 208-byte parameter block (src/shaders.mjs:9-24).
 */
#ifndef RAYMOND_H
#define RAYMOND_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RM_OK 0
#define RM_ERR_INVALID (-1)   /* a bad argument */
#define RM_ERR_DEVICE (-2)    /* a HIP runtime error */
#define RM_ERR_MEMORY (-3)    /* device memory exhausted */
#define RM_ERR_STATE (-4)     /* a call before what it needs (e.g. rm_frame before rm_resize) */
#define RM_ERR_CANCELLED (-5) /* a served job cancelled by rm_server_cancel, or past its deadline */

#define RM_PARAMS_BYTES 208

/* Buffers rm_read copies back, W·H·4 floats each. */
#define RM_BUFFER_ACCUM 0        /* per pixel Σ colour, Σ luminance² in w */
#define RM_BUFFER_GUIDE_ALBEDO 1 /* Σ first non-delta albedo */
#define RM_BUFFER_GUIDE_NORMAL 2 /* Σ first non-delta normal, Σ distance in w */
#define RM_BUFFER_DENOISED 3     /* the denoiser's output (an average) */

/* Stages timed by rm_stage_times, in this order. */
#define RM_STAGE_TRACE 0
#define RM_STAGE_LIGHT_TRACE 1
#define RM_STAGE_RESOLVE 2
#define RM_STAGE_GATHER 3
#define RM_STAGE_PREPARE 4
#define RM_STAGE_ATROUS 5
#define RM_STAGE_COMPOSE 6
#define RM_STAGE_CLEAR 7
#define RM_STAGE_RAYCAST 8
#define RM_STAGE_COUNT 9

/* Options for rm_set_option. */
#define RM_OPTION_SCHEDULE 0           /* how the path-tracing pass maps paths to lanes, RM_SCHEDULE_*;
                                          every schedule gives the same results, bit for bit */
#define RM_OPTION_PERSISTENT_BLOCKS 1  /* blocks of 256 threads for RM_SCHEDULE_PERSISTENT; 0 (default): fill the GPU */
#define RM_OPTION_PRECISION 2          /* RM_PRECISION_* flags; 0 (default): the plain FP32 path */
#define RM_OPTION_AUTOTUNE 3           /* job servers: frames per candidate with which rm_tune picks the schedule
                                          and waves for each new scene and image size; 0 (default): no tuning */
#define RM_OPTION_BATCH 4              /* job servers: 1 gives the frames of the jobs running at once one
                                          path-pass launch (rm_frame_batch); 0 (default): a launch per job */
#define RM_OPTION_WAVES 5              /* waves per SIMD that the RM_SCHEDULE_SPLIT and RM_SCHEDULE_PIXEL path
                                          passes are compiled for: 4 (or 0, the default) or 5; the same results
                                          bit for bit */
#define RM_OPTION_LIGHT_PATHS 6        /* light paths (and photon slots) a frame with light tracing; 0 (default):
                                          the reference's min(W·H·spp, 524288). A budget past that cap changes
                                          the image's light-traced part (more photons, less noise) */
#define RM_OPTION_STACK_DEPTH 7        /* RM_SCHEDULE_PIXEL at four waves: traversal stack entries kept in LDS,
                                          8 (or 0, the default) or 16; the same results bit for bit */
#define RM_OPTION_BLOCK_SHAPE 8        /* RM_SCHEDULE_PIXEL at four waves: RM_BLOCK_SQUARES (0, the default;
                                          16×16 tiles of 8×8 wave squares) or RM_BLOCK_ROWS (64×4 tiles, a wave
                                          a 64-pixel row); the same results bit for bit */
#define RM_OPTION_SPLIT_DEPTH 9        /* RM_SCHEDULE_SPLIT: bounces the head kernel runs before it queues the
                                          paths still alive, 1 to 8 (0, the default: 2) */
#define RM_OPTION_TUNE_CANDIDATES 10   /* rm_tune: bit k set tries candidate k (the order rm_tune lists them);
                                          0 (the default): all of them */
#define RM_OPTION_COUNT 11

#define RM_BLOCK_SQUARES 0
#define RM_BLOCK_ROWS 1

/* Reduced- and extended-precision variants (RM_OPTION_PRECISION flags). */
#define RM_PRECISION_QUANTIZED_BVH 1   /* traverse the 8-bit child-quantized tree (rm_bvh_quantize) in the
                                          per-pixel path pass, rm_raycast, light tracing and photon gather;
                                          decoded boxes enclose the FP32 ones, so no FP32 hit is missed */
#define RM_PRECISION_FP16_GUIDES 2     /* the denoiser's per-frame guide (normal, distance) in FP16; the guide
                                          sums stay FP32 and the passes compute in FP32 */
#define RM_PRECISION_COMPENSATED 4     /* per-pixel running sums (radiance, guide albedo, guide normal) are
                                          compensated FP32: each add's rounding error is kept and folded
                                          back for the denoiser and for rm_read / rm_display */

#define RM_SCHEDULE_PIXEL 0            /* one thread per pixel, each path run to its end */
#define RM_SCHEDULE_COMPACT 1          /* a block's live paths are packed into its first waves between bounces */
#define RM_SCHEDULE_PERSISTENT 2       /* persistent lanes that regenerate finished paths from a tile queue */
#define RM_SCHEDULE_WAVEFRONT 3        /* one kernel per stage (extend, shade, shadow, finish) over queues of
                                          live paths, the paths' state in device memory between stages */
#define RM_SCHEDULE_SORTED 4           /* as RM_SCHEDULE_COMPACT, and twice a bounce a block's live paths are
                                          sorted: by the kind of material hit after tracing, by ray-direction
                                          octant and specular-or-not after shading, so waves shade and trace
                                          alike paths */
#define RM_SCHEDULE_SPLIT 5            /* the schedule a context starts with: a thread a pixel for the first
                                          RM_OPTION_SPLIT_DEPTH bounces, then persistent waves over a queue of
                                          the paths still alive (frames of one sample per pixel in the colour
                                          view; RM_SCHEDULE_PIXEL runs the others) */

typedef struct rm_context rm_context;
typedef struct rm_scene rm_scene;

/* Opens HIP device `device` with its own stream.
 * Replaces Renderer.create's adapter/device request (src/Renderer.mjs:247-279). */
int rm_open(int device, rm_context** out);

/* Frees everything the context owns (its scenes included).
 * Replaces Renderer.dispose (src/Renderer.mjs:531-547). */
int rm_close(rm_context* ctx);

/* The message for the calling thread's last failure ("" if none). */
const char* rm_last_error(void);

/* Copies a packed scene to device memory.
 * primitive_slots counts every 64-byte record (the BVH's primitive_count,
 * then volumes); lights is the u32 table of light primitives, volume
 * primitives and the lights' cumulative chances as f32 bits.
 * Replaces Renderer.#upload's storage buffers (src/Renderer.mjs:550-626). */
int rm_scene_upload(rm_context* ctx, const void* primitives, uint32_t primitive_slots, uint32_t primitive_count,
                    const void* nodes, uint32_t node_count, const void* materials, uint32_t material_count,
                    const float* attributes, uint32_t attribute_vec4s, const uint32_t* lights, uint32_t lights_len,
                    rm_scene** out);

/* Releases a scene's device memory. */
int rm_scene_release(rm_context* ctx, rm_scene* scene);

/* RGBA8 texture layers (sRGB colour when normal_maps is 0, plain normal
 * maps when 1), width × height × layers × 4 bytes, for a scene's materials.
 * Replaces Renderer.#uploadTextures (src/Renderer.mjs:637-661). */
int rm_scene_textures(rm_context* ctx, rm_scene* scene, int normal_maps, const uint8_t* rgba, uint32_t width,
                      uint32_t height, uint32_t layers);

/* The environment map (RGBA half floats, width × height) and its importance
 * table (table_width × (table_height + 1) floats).
 * Replaces Renderer.#uploadEnvironment (src/Renderer.mjs:723-760). */
int rm_environment(rm_context* ctx, const uint16_t* rgba_half, uint32_t width, uint32_t height, const float* table,
                   uint32_t table_width, uint32_t table_height);

/* Builds a scene's BVH on the host: the tree BVH.build makes
 * (src/BVH.mjs:57-234) — binned SAH over 32 centroid bins, leaves of at most
 * max_leaf_size primitives unless splitting pays — node for node and bit for
 * bit, on `threads` threads (0: one per core). No device is needed.
 * mins/maxs: 3 doubles per primitive. nodes: room for max(1, 2·count − 1)
 * nodes of 32 bytes (f32 min[3], u32 first, f32 max[3], u32 count);
 * *node_count receives the number written. order: room for `count` u32;
 * order[i] is the original index of the i-th primitive in leaf order. */
int rm_bvh_build(const double* mins, const double* maxs, uint32_t count, uint32_t max_leaf_size, uint32_t threads,
                 void* nodes, uint32_t node_capacity, uint32_t* node_count, uint32_t* order);

/* The 8-bit child-quantized form of a packed tree (node_count nodes of 32
 * bytes), written to qnodes (node_count records of 32 bytes). An inner
 * node's record holds its FP32 min corner, a power-of-two step per axis and
 * both children's boxes as 8-bit offsets rounded outward, so a decoded box
 * always encloses the FP32 one; a leaf's holds its (first, count). No
 * counterpart in the reference, whose nodes are FP32 only (src/BVH.mjs:1-16
 * describes them); rm_scene_upload builds it for every scene, and
 * RM_PRECISION_QUANTIZED_BVH traverses it. Fails for boxes that aren't
 * finite or children outside their parent. No device is needed. */
int rm_bvh_quantize(const void* nodes, uint32_t node_count, void* qnodes);

/* The 4-wide quantized form of a packed tree: every other level folded
 * into its parent, so a wide node has two to four children, their boxes
 * quantized against the node's box as rm_bvh_quantize does, leaf ranges
 * inline; 64 bytes per wide node, written to `wide` (room for `capacity`;
 * node_count always suffices), *wide_count receiving how many (0 when the
 * root is a leaf). Fails, as rm_bvh_quantize does, for boxes that aren't
 * finite, and for trees where a walk could push more than the 64-entry stack
 * holds. No counterpart in the reference; rm_scene_upload builds it, and
 * RM_PRECISION_QUANTIZED_BVH walks it when it exists. No device is needed. */
int rm_bvh_quantize_wide(const void* nodes, uint32_t node_count, void* wide, uint32_t capacity,
                         uint32_t* wide_count);

/* Sizes the per-pixel buffers for width × height (reused while they fit).
 * Replaces Renderer.#resize (src/Renderer.mjs:766-799). */
int rm_resize(rm_context* ctx, uint32_t width, uint32_t height);

/* Enqueues one frame: the path-tracing pass, then light tracing, splat
 * resolve and photon gather when params.caustics is set (emitters must then
 * be given, emitter_vec4s of them), then the denoiser when params.denoise is
 * set. Returns once the work is queued.
 * Replaces the compute pass of Renderer.render (src/Renderer.mjs:441-489). */
int rm_frame(rm_context* ctx, const rm_scene* scene, const void* params, const float* emitters,
             uint32_t emitter_vec4s);

/* One frame on each of n contexts (at most 8, all on one device) with their
 * path passes in a single launch: each context's frame as rm_frame enqueues
 * it, except that the per-pixel path kernel runs once for all of them, the
 * grid's tiles split among the jobs, whatever their schedules. Every job
 * renders exactly as it would alone. The frames must walk the same kind of
 * tree (RM_PRECISION_QUANTIZED_BVH set alike). emitters and emitter_vec4s
 * may be NULL when no frame traces light. Arguments are all checked, and
 * every context's buffers allocated, before anything is queued, so a failure
 * (device memory running out) queues nothing. No counterpart in the reference, which
 * renders one image per Renderer; the job server uses it under
 * RM_OPTION_BATCH. */
int rm_frame_batch(rm_context* const* ctxs, const rm_scene* const* scenes, const void* const* params,
                   const float* const* emitters, const uint32_t* emitter_vec4s, uint32_t n);

/* Waits for the queued work and copies `floats` floats of a per-pixel buffer.
 * Replaces readPixels' copy and map (src/Renderer.mjs:502-528). */
int rm_read(rm_context* ctx, int buffer, float* dst, size_t floats);

/* Runs the display pass on the last frame and copies the image: W·H RGBA8
 * pixels, what render() draws to the canvas (the view's tone mapping, sRGB,
 * each channel rounded to 8 bits).
 * Replaces DISPLAY_SHADER's draw (src/shaders.mjs:2100-2152, src/Renderer.mjs:476-487). */
int rm_display(rm_context* ctx, uint8_t* rgba, size_t bytes);

/* Waits for all queued work (device.queue.onSubmittedWorkDone()). */
int rm_synchronize(rm_context* ctx);

/* Traces n rays (8 floats each: origin, t_max, direction, unused) and
 * writes 4 words per ray (t f32, primitive u32 or 0xffffffff, u f32, v f32);
 * any_hit stops at the first hit found. counters, if not NULL, receive 2
 * words per ray: node records and primitive records tested.
 * Replaces trace() as Scene.raycast mirrors it (src/shaders.mjs:389-447,
 * src/Scene.mjs:202-283). */
int rm_raycast(rm_context* ctx, const rm_scene* scene, const float* rays, uint32_t n, int any_hit, uint32_t* hits,
               uint32_t* counters);

/* rm_raycast for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * no counterpart in the reference. rays, hits and counters (NULL or 2 words
 * per ray) are DEVICE pointers in rm_raycast's layouts; the kernel is
 * enqueued on the context's stream and the call returns without waiting
 * (rm_synchronize waits). It allocates nothing: the traversal spill is a
 * static 28 MiB array each device holds until the process ends, so at most
 * RM_RAYCAST_STATIC_MAX_RAYS rays a call (more is RM_ERR_INVALID). Calls on
 * one device share that array: calls that overlap on one GPU (two contexts'
 * streams, or threads) race, and must be ordered by the caller. The same
 * kernel as rm_raycast, so the same hit records bit for bit. */
#define RM_RAYCAST_STATIC_MAX_RAYS 65536
int rm_raycast_static(rm_context* ctx, const rm_scene* scene, const float* rays, uint32_t n, int any_hit,
                      uint32_t* hits, uint32_t* counters);

/* Per-stage HIP event timing: enable or disable, read (waits; fills ms and
 * launch counts for RM_STAGE_COUNT stages, accumulated since the last
 * reset), and reset. No counterpart in the reference, which has no GPU timers. */
int rm_timing(rm_context* ctx, int enabled);
int rm_stage_times(rm_context* ctx, double* ms, uint32_t* launches, uint32_t n);
int rm_reset_stage_times(rm_context* ctx);

/* Sets an RM_OPTION_* for later frames. No counterpart in the reference:
 * these choose among equivalent kernel schedules (and let a tuner pick). A
 * context starts with RM_SCHEDULE_SPLIT and every other option at 0. */
int rm_set_option(rm_context* ctx, int option, int64_t value);

/* The autotuner. Renders `frames` whole frames to warm up, waits for the
 * GPU, then times `frames` more by events on the path stream (from a drained
 * start, as a caller that synchronises runs its frames), twice (the
 * candidates in turn, each keeping its faster time),
 * with the caller's parameters — light tracing (given its emitters) and the denoiser
 * included when they turn them on — under each candidate: split after two
 * bounces at four and at five waves per SIMD, and after one at five; one
 * thread per pixel at four and at five waves, at four with 16 stack entries
 * in LDS and with row-shaped blocks; compacted; sorted; persistent at two
 * grid sizes; wavefront (in that order of preference: a later candidate
 * must be at least 1% faster than the best before it). It sets the fastest as the context's
 * RM_OPTION_SCHEDULE, RM_OPTION_PERSISTENT_BLOCKS, RM_OPTION_WAVES,
 * RM_OPTION_STACK_DEPTH, RM_OPTION_BLOCK_SHAPE and RM_OPTION_SPLIT_DEPTH.
 * Whole frames, because the path pass shares the GPU with light tracing and
 * the previous frame's denoiser: a path pass that holds more waves can be
 * faster alone and still make the frame slower. The candidates give the same
 * sums bit for bit, so tuning changes speed only.
 * The frames render into the context's buffers: begin the image again
 * (sample offset 0) afterwards. On failure the context keeps the options it
 * had. No counterpart in the reference. */
#define RM_TUNE_CANDIDATES 12
typedef struct rm_tune_result {
    int32_t schedule[RM_TUNE_CANDIDATES];            /* each candidate's RM_SCHEDULE_* */
    uint32_t persistent_blocks[RM_TUNE_CANDIDATES];  /* and its persistent grid (0: fill the GPU) */
    uint32_t waves[RM_TUNE_CANDIDATES];              /* and its RM_OPTION_WAVES */
    uint32_t stack_depth[RM_TUNE_CANDIDATES];        /* and its RM_OPTION_STACK_DEPTH */
    uint32_t block_shape[RM_TUNE_CANDIDATES];        /* and its RM_OPTION_BLOCK_SHAPE */
    uint32_t split_depth[RM_TUNE_CANDIDATES];        /* and its RM_OPTION_SPLIT_DEPTH */
    double ms[RM_TUNE_CANDIDATES];                   /* mean ms per whole frame; 0: not tried (the
                                                        RM_OPTION_TUNE_CANDIDATES mask left it out) */
    uint32_t count;                                  /* candidates listed (RM_TUNE_CANDIDATES) */
    uint32_t best;                                   /* the one now set */
} rm_tune_result;
int rm_tune(rm_context* ctx, const rm_scene* scene, const void* params, const float* emitters,
            uint32_t emitter_vec4s, uint32_t frames, rm_tune_result* out);

/* The device's name and compute-unit count. */
int rm_device_info(rm_context* ctx, char* name, size_t name_len, int* compute_units);

/* ------------------------------------------------------------------ serving
 * A job server: render jobs (a scene, a parameter block, a target sample
 * count) arrive while others run. Up to `slots` jobs are active at once, each
 * in its own context with its own streams, so their frames overlap on the
 * device; more wait in a queue until a slot is free and the memory the job
 * needs fits in `memory_budget` bytes. Scenes stay resident: upload them on
 * rm_server_context(server) and refer to them from any number of jobs; release
 * a scene only after its jobs are done. A job renders frames exactly as
 * rm_frame does on a context of its own (sample offsets, running totals and
 * the photon radius advancing as the Renderer advances them), so its image is
 * the same whatever else is running. The denoiser, which reads only the
 * current accumulation, runs on a job's first frame and its last only. No
 * counterpart in the reference, which
 * renders one image per Renderer; this is the serving loop its callers would
 * otherwise write around Renderer.render. */
typedef struct rm_server rm_server;

typedef struct rm_job {
    const rm_scene* scene;         /* uploaded on rm_server_context() */
    const void* params;            /* the first frame's 208-byte block: size, samples per frame, features */
    const float* emitters;         /* light tracing's emitters when params.caustics is set, else NULL */
    uint32_t emitter_vec4s;
    const uint16_t* env_rgba;      /* the environment (as rm_environment takes it), or NULL for the sky */
    uint32_t env_width, env_height;
    const float* env_table;
    uint32_t table_width, table_height;
    uint32_t target_samples;       /* samples per pixel at which the job is done */
    double deadline_ms;            /* 0: none; else the job must have queued its last frame this many ms after
                                      submission, or it ends with RM_ERR_CANCELLED (as if cancelled then) */
} rm_job;

typedef struct rm_job_stats {
    double queued_ms;              /* submit → admitted to a slot */
    double first_frame_ms;         /* submit → its first frame (denoised if params.denoise) complete */
    double done_ms;                /* submit → target reached and the image copied out */
    uint32_t samples;              /* samples per pixel accumulated */
    uint32_t frames;
} rm_job_stats;

int rm_server_open(int device, uint32_t slots, uint64_t memory_budget, rm_server** out);
/* The context scenes are uploaded on (rm_scene_upload, rm_scene_textures). */
rm_context* rm_server_context(rm_server* server);
/* Copies the job's parameters, emitters and environment; returns its id. */
int rm_server_submit(rm_server* server, const rm_job* job, uint64_t* id);
/* Waits for a job and copies its image: the denoised buffer when the job's
 * params.denoise is set, else the accumulation (W·H·4 floats; dst may be NULL). */
int rm_server_wait(rm_server* server, uint64_t id, float* dst, size_t floats, rm_job_stats* stats);
/* Cancels a job: a queued job leaves the queue, a running one queues no
 * more frames and ends once those in flight are done; rm_server_wait then
 * returns RM_ERR_CANCELLED (its stats count the frames done). A job whose
 * last frame is already queued finishes with its image. */
int rm_server_cancel(rm_server* server, uint64_t id);
/* Sets an RM_OPTION_* (as rm_set_option does for a context) for the jobs
 * submitted from now on; each job keeps the options it was submitted with. */
int rm_server_set_option(rm_server* server, int option, int64_t value);
/* Waits for every job, then frees the server and its scenes. */
int rm_server_close(rm_server* server);

#ifdef __cplusplus
}
#endif
#endif /* RAYMOND_H */
