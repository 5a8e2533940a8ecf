/*
 * raymond.node — the thin N-API addon between the JS Renderer
 * (raymond_amd/js/renderer.mjs) and libraymond_hip.so. Each export maps one
 * C-ABI call; failures throw a JS Error with the library's message. Reads
 * and waits run on a worker thread and resolve a Promise, as WebGPU's
 * mapAsync and onSubmittedWorkDone do for the reference's Renderer.
 */
#include <node_api.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "raymond.h"

#define CHECK(call)                                                  \
    do {                                                             \
        if ((call) != napi_ok) {                                     \
            napi_throw_error(env, NULL, "N-API call failed: " #call); \
            return NULL;                                             \
        }                                                            \
    } while (0)

static napi_value throw_rm(napi_env env, int code) {
    char msg[512];
    snprintf(msg, sizeof msg, "%s (code %d)", rm_last_error(), code);
    napi_throw_error(env, NULL, msg);
    return NULL;
}

static int args(napi_env env, napi_callback_info info, size_t want, napi_value* argv) {
    size_t argc = want;
    if (napi_get_cb_info(env, info, &argc, argv, NULL, NULL) != napi_ok) return 0;
    if (argc < want) {
        napi_throw_type_error(env, NULL, "Too few arguments.");
        return 0;
    }
    return 1;
}

static int get_u32(napi_env env, napi_value v, uint32_t* out) {
    if (napi_get_value_uint32(env, v, out) != napi_ok) {
        napi_throw_type_error(env, NULL, "Expected a number.");
        return 0;
    }
    return 1;
}

static int get_ptr(napi_env env, napi_value v, void** out) {
    if (napi_get_value_external(env, v, out) != napi_ok || !*out) {
        napi_throw_type_error(env, NULL, "Expected a Raymond handle.");
        return 0;
    }
    return 1;
}

/* Data and byte length of an ArrayBuffer or typed array; null gives NULL. */
static int get_bytes(napi_env env, napi_value v, void** data, size_t* bytes) {
    napi_valuetype t;
    *data = NULL;
    *bytes = 0;
    if (napi_typeof(env, v, &t) != napi_ok) return 0;
    if (t == napi_null || t == napi_undefined) return 1;
    bool is_ab = false, is_ta = false;
    napi_is_arraybuffer(env, v, &is_ab);
    if (is_ab) return napi_get_arraybuffer_info(env, v, data, bytes) == napi_ok;
    napi_is_typedarray(env, v, &is_ta);
    if (is_ta) {
        napi_typedarray_type type;
        size_t length, offset;
        napi_value ab;
        void* base;
        if (napi_get_typedarray_info(env, v, &type, &length, &base, &ab, &offset) != napi_ok) return 0;
        size_t elem = 1;
        switch (type) {
        case napi_uint16_array: case napi_int16_array: elem = 2; break;
        case napi_uint32_array: case napi_int32_array: case napi_float32_array: elem = 4; break;
        case napi_float64_array: elem = 8; break;
        default: elem = 1;
        }
        *data = base;
        *bytes = length * elem;
        return 1;
    }
    napi_throw_type_error(env, NULL, "Expected an ArrayBuffer or typed array.");
    return 0;
}

static napi_value make_external(napi_env env, void* p) {
    napi_value out;
    CHECK(napi_create_external(env, p, NULL, NULL, &out));
    return out;
}

static napi_value undefined(napi_env env) {
    napi_value u;
    napi_get_undefined(env, &u);
    return u;
}

static napi_value js_open(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    uint32_t device = 0;
    size_t argc = 1;
    CHECK(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    if (argc >= 1 && !get_u32(env, argv[0], &device)) return NULL;
    rm_context* ctx = NULL;
    int r = rm_open((int)device, &ctx);
    if (r != RM_OK) return throw_rm(env, r);
    return make_external(env, ctx);
}

static napi_value js_close(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    void* ctx;
    if (!args(env, info, 1, argv) || !get_ptr(env, argv[0], &ctx)) return NULL;
    int r = rm_close((rm_context*)ctx);
    if (r != RM_OK) return throw_rm(env, r);
    return undefined(env);
}

static napi_value js_upload_scene(napi_env env, napi_callback_info info) {
    napi_value argv[9];
    void *ctx, *prims, *nodes, *mats, *attr, *lights;
    size_t pb, nb, mb, ab, lb;
    uint32_t prim_count, node_count, mat_count;
    if (!args(env, info, 9, argv) || !get_ptr(env, argv[0], &ctx)) return NULL;
    if (!get_bytes(env, argv[1], &prims, &pb) || !get_u32(env, argv[2], &prim_count) ||
        !get_bytes(env, argv[3], &nodes, &nb) || !get_u32(env, argv[4], &node_count) ||
        !get_bytes(env, argv[5], &mats, &mb) || !get_u32(env, argv[6], &mat_count) ||
        !get_bytes(env, argv[7], &attr, &ab) || !get_bytes(env, argv[8], &lights, &lb))
        return NULL;
    if (node_count * 32ull > nb || mat_count * 96ull > mb) {
        napi_throw_range_error(env, NULL, "A scene buffer is shorter than its count.");
        return NULL;
    }
    rm_scene* scene = NULL;
    int r = rm_scene_upload((rm_context*)ctx, prims, (uint32_t)(pb / 64), prim_count, nodes, node_count, mats,
                            mat_count, (const float*)attr, (uint32_t)(ab / 16), (const uint32_t*)lights,
                            (uint32_t)(lb / 4), &scene);
    if (r != RM_OK) return throw_rm(env, r);
    return make_external(env, scene);
}

static napi_value js_release_scene(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    void *ctx, *scene;
    if (!args(env, info, 2, argv) || !get_ptr(env, argv[0], &ctx) || !get_ptr(env, argv[1], &scene)) return NULL;
    int r = rm_scene_release((rm_context*)ctx, (rm_scene*)scene);
    if (r != RM_OK) return throw_rm(env, r);
    return undefined(env);
}

static napi_value js_upload_textures(napi_env env, napi_callback_info info) {
    napi_value argv[7];
    void *ctx, *scene, *data;
    size_t bytes;
    bool normal = false;
    uint32_t w, h, layers;
    if (!args(env, info, 7, argv) || !get_ptr(env, argv[0], &ctx) || !get_ptr(env, argv[1], &scene)) return NULL;
    napi_get_value_bool(env, argv[2], &normal);
    if (!get_bytes(env, argv[3], &data, &bytes) || !get_u32(env, argv[4], &w) || !get_u32(env, argv[5], &h) ||
        !get_u32(env, argv[6], &layers))
        return NULL;
    if ((uint64_t)w * h * layers * 4 > bytes) {
        napi_throw_range_error(env, NULL, "The texture data is shorter than its size.");
        return NULL;
    }
    int r = rm_scene_textures((rm_context*)ctx, (rm_scene*)scene, normal ? 1 : 0, (const uint8_t*)data, w, h, layers);
    if (r != RM_OK) return throw_rm(env, r);
    return undefined(env);
}

static napi_value js_upload_environment(napi_env env, napi_callback_info info) {
    napi_value argv[7];
    void *ctx, *rgba, *table;
    size_t rb, tb;
    uint32_t w, h, tw, th;
    if (!args(env, info, 7, argv) || !get_ptr(env, argv[0], &ctx)) return NULL;
    if (!get_bytes(env, argv[1], &rgba, &rb) || !get_u32(env, argv[2], &w) || !get_u32(env, argv[3], &h) ||
        !get_bytes(env, argv[4], &table, &tb) || !get_u32(env, argv[5], &tw) || !get_u32(env, argv[6], &th))
        return NULL;
    if ((uint64_t)w * h * 8 > rb || (uint64_t)tw * (th + 1) * 4 > tb) {
        napi_throw_range_error(env, NULL, "The environment data is shorter than its size.");
        return NULL;
    }
    int r = rm_environment((rm_context*)ctx, (const uint16_t*)rgba, w, h, (const float*)table, tw, th);
    if (r != RM_OK) return throw_rm(env, r);
    return undefined(env);
}

static napi_value js_resize(napi_env env, napi_callback_info info) {
    napi_value argv[3];
    void* ctx;
    uint32_t w, h;
    if (!args(env, info, 3, argv) || !get_ptr(env, argv[0], &ctx) || !get_u32(env, argv[1], &w) ||
        !get_u32(env, argv[2], &h))
        return NULL;
    int r = rm_resize((rm_context*)ctx, w, h);
    if (r != RM_OK) return throw_rm(env, r);
    return undefined(env);
}

static napi_value js_frame(napi_env env, napi_callback_info info) {
    napi_value argv[4];
    void *ctx, *scene, *params, *emitters;
    size_t pb, eb;
    if (!args(env, info, 4, argv) || !get_ptr(env, argv[0], &ctx) || !get_ptr(env, argv[1], &scene)) return NULL;
    if (!get_bytes(env, argv[2], &params, &pb) || !get_bytes(env, argv[3], &emitters, &eb)) return NULL;
    if (pb < RM_PARAMS_BYTES) {
        napi_throw_range_error(env, NULL, "The parameter block is 208 bytes.");
        return NULL;
    }
    int r = rm_frame((rm_context*)ctx, (const rm_scene*)scene, params, (const float*)emitters, (uint32_t)(eb / 16));
    if (r != RM_OK) return throw_rm(env, r);
    return undefined(env);
}

/* tune(ctx, scene, params, emitters, frames): rm_tune, which renders whole
 * frames under each candidate and keeps the fastest on the context. Blocks
 * until the frames are done. Returns {best, candidates: [{schedule,
 * persistentBlocks, waves, stackDepth, blockShape, splitDepth, ms}]}, best
 * an index into candidates. */
static napi_value candidate_object(napi_env env, const rm_tune_result* res, uint32_t k) {
    napi_value o, v;
    CHECK(napi_create_object(env, &o));
    CHECK(napi_create_int32(env, res->schedule[k], &v));
    CHECK(napi_set_named_property(env, o, "schedule", v));
    CHECK(napi_create_uint32(env, res->persistent_blocks[k], &v));
    CHECK(napi_set_named_property(env, o, "persistentBlocks", v));
    CHECK(napi_create_uint32(env, res->waves[k], &v));
    CHECK(napi_set_named_property(env, o, "waves", v));
    CHECK(napi_create_uint32(env, res->stack_depth[k], &v));
    CHECK(napi_set_named_property(env, o, "stackDepth", v));
    CHECK(napi_create_uint32(env, res->block_shape[k], &v));
    CHECK(napi_set_named_property(env, o, "blockShape", v));
    CHECK(napi_create_uint32(env, res->split_depth[k], &v));
    CHECK(napi_set_named_property(env, o, "splitDepth", v));
    CHECK(napi_create_double(env, res->ms[k], &v));
    CHECK(napi_set_named_property(env, o, "ms", v));
    return o;
}

static napi_value js_tune(napi_env env, napi_callback_info info) {
    napi_value argv[5];
    void *ctx, *scene, *params, *emitters;
    size_t pb, eb;
    uint32_t frames;
    if (!args(env, info, 5, argv) || !get_ptr(env, argv[0], &ctx) || !get_ptr(env, argv[1], &scene)) return NULL;
    if (!get_bytes(env, argv[2], &params, &pb) || !get_bytes(env, argv[3], &emitters, &eb)) return NULL;
    if (!get_u32(env, argv[4], &frames)) return NULL;
    if (pb < RM_PARAMS_BYTES) {
        napi_throw_range_error(env, NULL, "The parameter block is 208 bytes.");
        return NULL;
    }
    rm_tune_result res;
    int r = rm_tune((rm_context*)ctx, (const rm_scene*)scene, params, (const float*)emitters, (uint32_t)(eb / 16),
                    frames, &res);
    if (r != RM_OK) return throw_rm(env, r);
    napi_value out, list, v;
    CHECK(napi_create_object(env, &out));
    CHECK(napi_create_array_with_length(env, res.count, &list));
    for (uint32_t k = 0; k < res.count; ++k) {
        napi_value c = candidate_object(env, &res, k);
        if (!c) return NULL;
        CHECK(napi_set_element(env, list, k, c));
    }
    CHECK(napi_set_named_property(env, out, "candidates", list));
    CHECK(napi_create_uint32(env, res.best, &v));
    CHECK(napi_set_named_property(env, out, "best", v));
    return out;
}

/* ---- asynchronous reads and waits ---- */

typedef struct {
    napi_async_work work;
    napi_deferred deferred;
    napi_ref keep;          /* the destination array, alive until done */
    rm_context* ctx;
    int buffer;             /* -1: synchronize only; -2: the display pass into `dst` as bytes */
    float* dst;
    size_t floats;          /* for the display pass: bytes */
    rm_server* server;      /* non-null: wait for job `job` of this server */
    uint64_t job;
    rm_job_stats stats;
    int result;
    char message[512];
} AsyncJob;

static void job_execute(napi_env env, void* data) {
    (void)env;
    AsyncJob* j = (AsyncJob*)data;
    if (j->server)
        j->result = rm_server_wait(j->server, j->job, j->dst, j->floats, &j->stats);
    else if (j->buffer == -2)
        j->result = rm_display(j->ctx, (uint8_t*)j->dst, j->floats);
    else
        j->result = j->buffer < 0 ? rm_synchronize(j->ctx) : rm_read(j->ctx, j->buffer, j->dst, j->floats);
    if (j->result != RM_OK) snprintf(j->message, sizeof j->message, "%s (code %d)", rm_last_error(), j->result);
}

static napi_value stats_object(napi_env env, const rm_job_stats* st) {
    napi_value o, v;
    napi_create_object(env, &o);
    napi_create_double(env, st->queued_ms, &v);
    napi_set_named_property(env, o, "queuedMs", v);
    napi_create_double(env, st->first_frame_ms, &v);
    napi_set_named_property(env, o, "firstFrameMs", v);
    napi_create_double(env, st->done_ms, &v);
    napi_set_named_property(env, o, "doneMs", v);
    napi_create_uint32(env, st->samples, &v);
    napi_set_named_property(env, o, "samples", v);
    napi_create_uint32(env, st->frames, &v);
    napi_set_named_property(env, o, "frames", v);
    return o;
}

static void job_complete(napi_env env, napi_status status, void* data) {
    AsyncJob* j = (AsyncJob*)data;
    if (status == napi_ok && j->result == RM_OK) {
        napi_resolve_deferred(env, j->deferred, j->server ? stats_object(env, &j->stats) : undefined(env));
    } else {
        napi_value msg, err;
        napi_create_string_utf8(env, j->result != RM_OK ? j->message : "The GPU wait was cancelled.", NAPI_AUTO_LENGTH,
                                &msg);
        napi_create_error(env, NULL, msg, &err);
        napi_reject_deferred(env, j->deferred, err);
    }
    if (j->keep) napi_delete_reference(env, j->keep);
    napi_delete_async_work(env, j->work);
    free(j);
}

static napi_value start_job(napi_env env, AsyncJob* j, napi_value keep) {
    napi_value promise, name;
    if (napi_create_promise(env, &j->deferred, &promise) != napi_ok) {
        free(j);
        napi_throw_error(env, NULL, "Couldn't create a promise.");
        return NULL;
    }
    if (keep) napi_create_reference(env, keep, 1, &j->keep);
    napi_create_string_utf8(env, "raymond", NAPI_AUTO_LENGTH, &name);
    napi_create_async_work(env, NULL, name, job_execute, job_complete, j, &j->work);
    napi_queue_async_work(env, j->work);
    return promise;
}

static napi_value js_read(napi_env env, napi_callback_info info) {
    napi_value argv[3];
    void *ctx, *dst;
    size_t bytes;
    uint32_t which;
    if (!args(env, info, 3, argv) || !get_ptr(env, argv[0], &ctx) || !get_u32(env, argv[1], &which) ||
        !get_bytes(env, argv[2], &dst, &bytes))
        return NULL;
    AsyncJob* j = (AsyncJob*)calloc(1, sizeof *j);
    j->ctx = (rm_context*)ctx;
    j->buffer = (int)which;
    j->dst = (float*)dst;
    j->floats = bytes / 4;
    return start_job(env, j, argv[2]);
}

/* display(ctx, Uint8Array of W·H·4) → Promise: the display pass on the last frame. */
static napi_value js_display(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    void *ctx, *dst;
    size_t bytes;
    if (!args(env, info, 2, argv) || !get_ptr(env, argv[0], &ctx) || !get_bytes(env, argv[1], &dst, &bytes)) return NULL;
    AsyncJob* j = (AsyncJob*)calloc(1, sizeof *j);
    j->ctx = (rm_context*)ctx;
    j->buffer = -2;
    j->dst = (float*)dst;
    j->floats = bytes;
    return start_job(env, j, argv[1]);
}

static napi_value js_synchronize(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    void* ctx;
    if (!args(env, info, 1, argv) || !get_ptr(env, argv[0], &ctx)) return NULL;
    AsyncJob* j = (AsyncJob*)calloc(1, sizeof *j);
    j->ctx = (rm_context*)ctx;
    j->buffer = -1;
    return start_job(env, j, NULL);
}

static napi_value js_raycast(napi_env env, napi_callback_info info) {
    napi_value argv[4];
    void *ctx, *scene, *rays;
    size_t bytes;
    bool any = false;
    if (!args(env, info, 4, argv) || !get_ptr(env, argv[0], &ctx) || !get_ptr(env, argv[1], &scene) ||
        !get_bytes(env, argv[2], &rays, &bytes))
        return NULL;
    napi_get_value_bool(env, argv[3], &any);
    uint32_t n = (uint32_t)(bytes / 32);
    void* out_data;
    napi_value ab, hits;
    CHECK(napi_create_arraybuffer(env, (size_t)n * 16 + 16, &out_data, &ab));
    CHECK(napi_create_typedarray(env, napi_uint32_array, (size_t)n * 4, ab, 0, &hits));
    int r = rm_raycast((rm_context*)ctx, (const rm_scene*)scene, (const float*)rays, n, any ? 1 : 0,
                       (uint32_t*)out_data, NULL);
    if (r != RM_OK) return throw_rm(env, r);
    return hits;
}

static napi_value js_device_info(napi_env env, napi_callback_info info) {
    napi_value argv[1], obj, name, cus;
    void* ctx;
    char buf[256];
    int n = 0;
    if (!args(env, info, 1, argv) || !get_ptr(env, argv[0], &ctx)) return NULL;
    int r = rm_device_info((rm_context*)ctx, buf, sizeof buf, &n);
    if (r != RM_OK) return throw_rm(env, r);
    CHECK(napi_create_object(env, &obj));
    CHECK(napi_create_string_utf8(env, buf, NAPI_AUTO_LENGTH, &name));
    CHECK(napi_create_int32(env, n, &cus));
    CHECK(napi_set_named_property(env, obj, "name", name));
    CHECK(napi_set_named_property(env, obj, "computeUnits", cus));
    return obj;
}

/* A Float64Array's data and length (elements). */
static int get_f64(napi_env env, napi_value v, const double** data, size_t* length) {
    bool is_ta = false;
    napi_typedarray_type type;
    size_t offset;
    napi_value ab;
    void* base;
    napi_is_typedarray(env, v, &is_ta);
    if (!is_ta || napi_get_typedarray_info(env, v, &type, length, &base, &ab, &offset) != napi_ok ||
        type != napi_float64_array) {
        napi_throw_type_error(env, NULL, "Expected a Float64Array.");
        return 0;
    }
    *data = (const double*)base;
    return 1;
}

/* bvhBuild(mins, maxs, maxLeafSize, threads) → { nodes, nodeCount, order }:
 * BVH.build's result (src/BVH.mjs:57-234) from the native builder. */
static napi_value js_bvh_build(napi_env env, napi_callback_info info) {
    napi_value argv[4], obj, nodes_ab, order_ab, order, count_v;
    const double *mins, *maxs;
    size_t n_min, n_max;
    uint32_t max_leaf, threads;
    if (!args(env, info, 4, argv) || !get_f64(env, argv[0], &mins, &n_min) || !get_f64(env, argv[1], &maxs, &n_max) ||
        !get_u32(env, argv[2], &max_leaf) || !get_u32(env, argv[3], &threads))
        return NULL;
    if (n_min != n_max || n_min % 3 != 0 || n_min / 3 >= (1u << 31)) {
        napi_throw_range_error(env, NULL, "mins and maxs need 3 values per primitive, as many of each.");
        return NULL;
    }
    uint32_t count = (uint32_t)(n_min / 3), most = count > 1 ? 2 * count - 1 : 1, used = 0;
    void* order_data;
    CHECK(napi_create_arraybuffer(env, (size_t)count * 4, &order_data, &order_ab));
    CHECK(napi_create_typedarray(env, napi_uint32_array, count, order_ab, 0, &order));
    void* scratch = malloc((size_t)most * 32);
    if (!scratch) {
        napi_throw_error(env, NULL, "Out of memory for the BVH's nodes.");
        return NULL;
    }
    int r = rm_bvh_build(mins, maxs, count, max_leaf, threads, scratch, most, &used, (uint32_t*)order_data);
    if (r != RM_OK) {
        free(scratch);
        return throw_rm(env, r);
    }
    void* nodes_data;
    if (napi_create_arraybuffer(env, (size_t)used * 32, &nodes_data, &nodes_ab) != napi_ok) {
        free(scratch);
        napi_throw_error(env, NULL, "N-API call failed: napi_create_arraybuffer");
        return NULL;
    }
    memcpy(nodes_data, scratch, (size_t)used * 32);
    free(scratch);
    CHECK(napi_create_object(env, &obj));
    CHECK(napi_create_uint32(env, used, &count_v));
    CHECK(napi_set_named_property(env, obj, "nodes", nodes_ab));
    CHECK(napi_set_named_property(env, obj, "nodeCount", count_v));
    CHECK(napi_set_named_property(env, obj, "order", order));
    return obj;
}

/* ---- the job server ---- */

static napi_value js_server_open(napi_env env, napi_callback_info info) {
    napi_value argv[3];
    uint32_t device, slots;
    double budget;
    if (!args(env, info, 3, argv) || !get_u32(env, argv[0], &device) || !get_u32(env, argv[1], &slots)) return NULL;
    if (napi_get_value_double(env, argv[2], &budget) != napi_ok || !(budget >= 0.0)) {
        napi_throw_type_error(env, NULL, "Expected a memory budget in bytes.");
        return NULL;
    }
    rm_server* sv = NULL;
    int r = rm_server_open((int)device, slots, (uint64_t)budget, &sv);
    if (r != RM_OK) return throw_rm(env, r);
    return make_external(env, sv);
}

static napi_value js_server_context(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    void* sv;
    if (!args(env, info, 1, argv) || !get_ptr(env, argv[0], &sv)) return NULL;
    return make_external(env, rm_server_context((rm_server*)sv));
}

/* serverSubmit(server, scene, params, targetSamples, emitters|null, envRgba|null, envWidth, envHeight,
 *              envTable|null, tableWidth, tableHeight, deadlineMs) → job id (a Number) */
static napi_value js_server_submit(napi_env env, napi_callback_info info) {
    napi_value argv[12];
    void *sv, *scene, *params, *emitters, *rgba, *table;
    size_t pb, eb, rb, tb;
    uint32_t target, ew, eh, tw, th;
    double deadline = 0.0;
    if (!args(env, info, 12, argv) || !get_ptr(env, argv[0], &sv) || !get_ptr(env, argv[1], &scene) ||
        !get_bytes(env, argv[2], &params, &pb) || !get_u32(env, argv[3], &target) ||
        !get_bytes(env, argv[4], &emitters, &eb) || !get_bytes(env, argv[5], &rgba, &rb) ||
        !get_u32(env, argv[6], &ew) || !get_u32(env, argv[7], &eh) || !get_bytes(env, argv[8], &table, &tb) ||
        !get_u32(env, argv[9], &tw) || !get_u32(env, argv[10], &th))
        return NULL;
    if (napi_get_value_double(env, argv[11], &deadline) != napi_ok) {
        napi_throw_type_error(env, NULL, "Expected a deadline in milliseconds.");
        return NULL;
    }
    if (pb < RM_PARAMS_BYTES) {
        napi_throw_range_error(env, NULL, "The parameter block is shorter than 208 bytes.");
        return NULL;
    }
    rm_job job;
    memset(&job, 0, sizeof job);
    job.scene = (const rm_scene*)scene;
    job.params = params;
    job.emitters = (const float*)emitters;
    job.emitter_vec4s = (uint32_t)(eb / 16);
    if (rgba) {
        job.env_rgba = (const uint16_t*)rgba;
        job.env_width = ew;
        job.env_height = eh;
        job.env_table = (const float*)table;
        job.table_width = tw;
        job.table_height = th;
    }
    job.target_samples = target;
    job.deadline_ms = deadline;
    uint64_t id = 0;
    int r = rm_server_submit((rm_server*)sv, &job, &id);
    if (r != RM_OK) return throw_rm(env, r);
    napi_value out;
    CHECK(napi_create_double(env, (double)id, &out));
    return out;
}

/* serverCancel(server, id) */
static napi_value js_server_cancel(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    void* sv;
    double id;
    if (!args(env, info, 2, argv) || !get_ptr(env, argv[0], &sv)) return NULL;
    if (napi_get_value_double(env, argv[1], &id) != napi_ok) {
        napi_throw_type_error(env, NULL, "Expected a job id.");
        return NULL;
    }
    int r = rm_server_cancel((rm_server*)sv, (uint64_t)id);
    if (r != RM_OK) return throw_rm(env, r);
    return undefined(env);
}

/* serverWait(server, id, Float32Array|null) → Promise of the job's timings */
static napi_value js_server_wait(napi_env env, napi_callback_info info) {
    napi_value argv[3];
    void *sv, *dst;
    size_t bytes;
    double id;
    if (!args(env, info, 3, argv) || !get_ptr(env, argv[0], &sv) || !get_bytes(env, argv[2], &dst, &bytes)) return NULL;
    if (napi_get_value_double(env, argv[1], &id) != napi_ok) {
        napi_throw_type_error(env, NULL, "Expected a job id.");
        return NULL;
    }
    AsyncJob* j = (AsyncJob*)calloc(1, sizeof *j);
    j->server = (rm_server*)sv;
    j->job = (uint64_t)id;
    j->dst = (float*)dst;
    j->floats = bytes / 4;
    return start_job(env, j, dst ? argv[2] : NULL);
}

static napi_value js_server_close(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    void* sv;
    if (!args(env, info, 1, argv) || !get_ptr(env, argv[0], &sv)) return NULL;
    int r = rm_server_close((rm_server*)sv);
    if (r != RM_OK) return throw_rm(env, r);
    return undefined(env);
}

/* setOption(ctx, option, value) and serverSetOption(server, option, value):
 * rm_set_option / rm_server_set_option. */
static napi_value set_option(napi_env env, napi_callback_info info, int server) {
    napi_value argv[3];
    void* h;
    uint32_t option;
    int64_t value;
    if (!args(env, info, 3, argv) || !get_ptr(env, argv[0], &h) || !get_u32(env, argv[1], &option)) return NULL;
    if (napi_get_value_int64(env, argv[2], &value) != napi_ok) {
        napi_throw_type_error(env, NULL, "Expected a number.");
        return NULL;
    }
    int r = server ? rm_server_set_option((rm_server*)h, (int)option, value)
                   : rm_set_option((rm_context*)h, (int)option, value);
    if (r != RM_OK) return throw_rm(env, r);
    return undefined(env);
}
static napi_value js_set_option(napi_env env, napi_callback_info info) { return set_option(env, info, 0); }
static napi_value js_server_set_option(napi_env env, napi_callback_info info) { return set_option(env, info, 1); }

static napi_value init(napi_env env, napi_value exports) {
    napi_property_descriptor props[] = {
        {"open", NULL, js_open, NULL, NULL, NULL, napi_enumerable, NULL},
        {"close", NULL, js_close, NULL, NULL, NULL, napi_enumerable, NULL},
        {"uploadScene", NULL, js_upload_scene, NULL, NULL, NULL, napi_enumerable, NULL},
        {"releaseScene", NULL, js_release_scene, NULL, NULL, NULL, napi_enumerable, NULL},
        {"uploadTextures", NULL, js_upload_textures, NULL, NULL, NULL, napi_enumerable, NULL},
        {"uploadEnvironment", NULL, js_upload_environment, NULL, NULL, NULL, napi_enumerable, NULL},
        {"resize", NULL, js_resize, NULL, NULL, NULL, napi_enumerable, NULL},
        {"frame", NULL, js_frame, NULL, NULL, NULL, napi_enumerable, NULL},
        {"read", NULL, js_read, NULL, NULL, NULL, napi_enumerable, NULL},
        {"synchronize", NULL, js_synchronize, NULL, NULL, NULL, napi_enumerable, NULL},
        {"display", NULL, js_display, NULL, NULL, NULL, napi_enumerable, NULL},
        {"raycast", NULL, js_raycast, NULL, NULL, NULL, napi_enumerable, NULL},
        {"deviceInfo", NULL, js_device_info, NULL, NULL, NULL, napi_enumerable, NULL},
        {"bvhBuild", NULL, js_bvh_build, NULL, NULL, NULL, napi_enumerable, NULL},
        {"setOption", NULL, js_set_option, NULL, NULL, NULL, napi_enumerable, NULL},
        {"tune", NULL, js_tune, NULL, NULL, NULL, napi_enumerable, NULL},
        {"serverSetOption", NULL, js_server_set_option, NULL, NULL, NULL, napi_enumerable, NULL},
        {"serverOpen", NULL, js_server_open, NULL, NULL, NULL, napi_enumerable, NULL},
        {"serverContext", NULL, js_server_context, NULL, NULL, NULL, napi_enumerable, NULL},
        {"serverSubmit", NULL, js_server_submit, NULL, NULL, NULL, napi_enumerable, NULL},
        {"serverWait", NULL, js_server_wait, NULL, NULL, NULL, napi_enumerable, NULL},
        {"serverCancel", NULL, js_server_cancel, NULL, NULL, NULL, napi_enumerable, NULL},
        {"serverClose", NULL, js_server_close, NULL, NULL, NULL, napi_enumerable, NULL},
    };
    if (napi_define_properties(env, exports, sizeof props / sizeof props[0], props) != napi_ok) return NULL;
    return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
