"""ctypes binding of libraymond_hip.so (include/raymond.h) for Python harnesses.

The JS Renderer reaches the same library through raymond.node; this module
is how the Python tests and bench drive it. There is no fallback: if the
library is missing or no GPU is present, opening a context raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libraymond_hip.so"

RM_OK = 0
ERR_INVALID, ERR_DEVICE, ERR_MEMORY, ERR_STATE, ERR_CANCELLED = -1, -2, -3, -4, -5
BUFFER_ACCUM, BUFFER_GUIDE_ALBEDO, BUFFER_GUIDE_NORMAL, BUFFER_DENOISED = 0, 1, 2, 3
STAGES = ("trace", "light_trace", "resolve", "gather", "prepare", "atrous", "compose", "clear", "raycast")

# Every function include/raymond.h declares, with its ctypes signature.
_P = C.c_void_p
_U32 = C.c_uint32
SIGNATURES = {
    "rm_open": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "rm_close": (C.c_int, [_P]),
    "rm_last_error": (C.c_char_p, []),
    "rm_scene_upload": (C.c_int, [_P, _P, _U32, _U32, _P, _U32, _P, _U32, _P, _U32, _P, _U32, C.POINTER(_P)]),
    "rm_scene_release": (C.c_int, [_P, _P]),
    "rm_scene_textures": (C.c_int, [_P, _P, C.c_int, _P, _U32, _U32, _U32]),
    "rm_environment": (C.c_int, [_P, _P, _U32, _U32, _P, _U32, _U32]),
    "rm_resize": (C.c_int, [_P, _U32, _U32]),
    "rm_frame_batch": (C.c_int, [_P, _P, _P, _P, _P, C.c_uint32]),
    "rm_frame": (C.c_int, [_P, _P, _P, _P, _U32]),
    "rm_read": (C.c_int, [_P, C.c_int, _P, C.c_size_t]),
    "rm_synchronize": (C.c_int, [_P]),
    "rm_display": (C.c_int, [_P, _P, C.c_size_t]),
    "rm_bvh_build": (C.c_int, [_P, _P, C.c_uint32, C.c_uint32, C.c_uint32, _P, C.c_uint32, _P, _P]),
    "rm_bvh_quantize": (C.c_int, [_P, C.c_uint32, _P]),
    "rm_bvh_quantize_wide": (C.c_int, [_P, C.c_uint32, _P, C.c_uint32, _P]),
    "rm_raycast": (C.c_int, [_P, _P, _P, _U32, C.c_int, _P, _P]),
    "rm_raycast_static": (C.c_int, [_P, _P, _P, _U32, C.c_int, _P, _P]),
    "rm_timing": (C.c_int, [_P, C.c_int]),
    "rm_stage_times": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(_U32), _U32]),
    "rm_reset_stage_times": (C.c_int, [_P]),
    "rm_device_info": (C.c_int, [_P, C.c_char_p, C.c_size_t, C.POINTER(C.c_int)]),
    "rm_tune": (C.c_int, [_P, _P, _P, _P, C.c_uint32, C.c_uint32, _P]),
    "rm_set_option": (C.c_int, [_P, C.c_int, C.c_int64]),
    "rm_server_open": (C.c_int, [C.c_int, _U32, C.c_uint64, C.POINTER(_P)]),
    "rm_server_context": (_P, [_P]),
    "rm_server_submit": (C.c_int, [_P, _P, C.POINTER(C.c_uint64)]),
    "rm_server_wait": (C.c_int, [_P, C.c_uint64, _P, C.c_size_t, _P]),
    "rm_server_close": (C.c_int, [_P]),
    "rm_server_cancel": (C.c_int, [_P, C.c_uint64]),
    "rm_server_set_option": (C.c_int, [_P, C.c_int, C.c_int64]),
}
(OPTION_SCHEDULE, OPTION_PERSISTENT_BLOCKS, OPTION_PRECISION, OPTION_AUTOTUNE, OPTION_BATCH, OPTION_WAVES,
 OPTION_LIGHT_PATHS, OPTION_STACK_DEPTH, OPTION_BLOCK_SHAPE, OPTION_SPLIT_DEPTH, OPTION_TUNE_CANDIDATES) = range(11)
BLOCK_SQUARES, BLOCK_ROWS = 0, 1
PRECISION_QUANTIZED_BVH, PRECISION_FP16_GUIDES, PRECISION_COMPENSATED = 1, 2, 4
SCHEDULE_PIXEL, SCHEDULE_COMPACT, SCHEDULE_PERSISTENT, SCHEDULE_WAVEFRONT, SCHEDULE_SORTED, SCHEDULE_SPLIT = range(6)
SCHEDULES = {"per-pixel": SCHEDULE_PIXEL, "compact": SCHEDULE_COMPACT, "persistent": SCHEDULE_PERSISTENT,
             "wavefront": SCHEDULE_WAVEFRONT, "sorted": SCHEDULE_SORTED, "split": SCHEDULE_SPLIT}
SCHEDULE_DEFAULT = SCHEDULE_SPLIT   # the schedule a context (and a server's jobs) start with

_lib = None


TUNE_CANDIDATES = 12


class _TuneResult(C.Structure):
    _fields_ = [("schedule", C.c_int32 * TUNE_CANDIDATES), ("persistent_blocks", C.c_uint32 * TUNE_CANDIDATES),
                ("waves", C.c_uint32 * TUNE_CANDIDATES), ("stack_depth", C.c_uint32 * TUNE_CANDIDATES),
                ("block_shape", C.c_uint32 * TUNE_CANDIDATES), ("split_depth", C.c_uint32 * TUNE_CANDIDATES),
                ("ms", C.c_double * TUNE_CANDIDATES),
                ("count", C.c_uint32), ("best", C.c_uint32)]


class RaymondError(RuntimeError):
    """A failed call; `code` is its RM_ERR_* value."""

    def __init__(self, message: str, code: int = 0):
        super().__init__(message)
        self.code = code


def load() -> C.CDLL:
    """Loads the library, failing loudly when it isn't built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RaymondError(f"{LIB_PATH} is not built: run python -m raymond_amd.build")
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, argt) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = argt
        _lib = lib
    return _lib


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(code: int, what: str) -> None:
    if code != RM_OK:
        raise RaymondError(f"{what}: {load().rm_last_error().decode()} (code {code})", code)


def frame_batch(contexts, scenes, params, emitters=None) -> None:
    """rm_frame_batch: one frame on each context, their path passes in one launch."""
    n = len(contexts)
    ctxs = (C.c_void_p * n)(*[c.handle.value if isinstance(c.handle, C.c_void_p) else c.handle for c in contexts])
    scs = (C.c_void_p * n)(*[s.handle.value if isinstance(s.handle, C.c_void_p) else s.handle for s in scenes])
    blocks = [np.ascontiguousarray(p) for p in params]
    pp = (C.c_void_p * n)(*[b.ctypes.data for b in blocks])
    em = ev = None
    if emitters is not None and any(e is not None for e in emitters):
        es = [None if e is None else np.ascontiguousarray(e, np.float32) for e in emitters]
        em = (C.c_void_p * n)(*[None if e is None else e.ctypes.data for e in es])
        ev = (C.c_uint32 * n)(*[0 if e is None else e.size // 4 for e in es])
    _check(load().rm_frame_batch(ctxs, scs, pp, em, ev, n), "rm_frame_batch")


def bvh_build(mins: np.ndarray, maxs: np.ndarray, max_leaf_size: int = 4, threads: int = 0):
    """The scene BVH BVH.build makes (src/BVH.mjs:57-234), built natively on
    the host: (nodes as (N, 8) u32 — f32 min[3], first, f32 max[3], count —
    and the leaf order)."""
    lo = np.ascontiguousarray(mins, dtype=np.float64).reshape(-1)
    hi = np.ascontiguousarray(maxs, dtype=np.float64).reshape(-1)
    if lo.size != hi.size or lo.size % 3:
        raise ValueError("mins and maxs need 3 values per primitive, as many of each")
    n = lo.size // 3
    most = max(1, 2 * n - 1)
    nodes = np.empty((most, 8), dtype=np.uint32)
    order = np.empty(n, dtype=np.uint32)
    used = C.c_uint32(0)
    _check(load().rm_bvh_build(_ptr(lo), _ptr(hi), n, max_leaf_size, threads, _ptr(nodes), most, C.byref(used),
                               _ptr(order)), "rm_bvh_build")
    return nodes[:used.value].copy(), order


def bvh_quantize(nodes: np.ndarray) -> np.ndarray:
    """The 8-bit child-quantized records of a packed tree ((N, 8) u32 in, (N, 8) u32 out)."""
    n = np.ascontiguousarray(nodes, dtype=np.uint32).reshape(-1, 8)
    out = np.empty_like(n)
    _check(load().rm_bvh_quantize(_ptr(n), n.shape[0], _ptr(out)), "rm_bvh_quantize")
    return out


def bvh_quantize_wide(nodes: np.ndarray) -> np.ndarray:
    """The 4-wide quantized records of a packed tree ((N, 8) u32 in, (W, 16) u32 out)."""
    n = np.ascontiguousarray(nodes, dtype=np.uint32).reshape(-1, 8)
    out = np.zeros((n.shape[0], 16), np.uint32)
    used = C.c_uint32(0)
    _check(load().rm_bvh_quantize_wide(_ptr(n), n.shape[0], _ptr(out), n.shape[0], C.byref(used)),
           "rm_bvh_quantize_wide")
    return out[:used.value].copy()


class Scene:
    def __init__(self, ctx: "Context", handle: C.c_void_p):
        self.ctx = ctx
        self.handle = handle

    def release(self) -> None:
        if self.handle:
            _check(load().rm_scene_release(self.ctx.handle, self.handle), "rm_scene_release")
            self.handle = None


class Context:
    """One device context: a stream, per-pixel buffers and uploaded scenes."""

    def __init__(self, device: int = 0, handle=None):
        """Opens device `device`, or wraps an existing handle (not closed by close())."""
        self.owned = handle is None
        if handle is None:
            lib = load()
            handle = C.c_void_p()
            _check(lib.rm_open(device, C.byref(handle)), "rm_open")
        self.handle = handle
        self.width = self.height = 0

    def close(self) -> None:
        if self.handle and self.owned:
            _check(load().rm_close(self.handle), "rm_close")
        self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload_scene(self, packed) -> Scene:
        """packed: a raymond_amd.scenefile.Packed (or anything with the same arrays)."""
        lib = load()
        h = C.c_void_p()
        prims = np.ascontiguousarray(packed.primitives, dtype=np.uint32)
        nodes = np.ascontiguousarray(packed.nodes, dtype=np.uint32)
        mats = np.ascontiguousarray(packed.materials, dtype=np.uint32)
        attr = np.ascontiguousarray(packed.attributes, dtype=np.float32)
        lights = np.ascontiguousarray(packed.lights, dtype=np.uint32)
        _check(lib.rm_scene_upload(self.handle, _ptr(prims), prims.size // 16, packed.primitive_count,
                                   _ptr(nodes), nodes.size // 8, _ptr(mats), mats.size // 24,
                                   _ptr(attr), attr.size // 4, _ptr(lights), lights.size, C.byref(h)),
               "rm_scene_upload")
        scene = Scene(self, h)
        for layers, normal in ((getattr(packed, "textures", None), False), (getattr(packed, "normal_maps", None), True)):
            if layers is not None:
                self.textures(scene, normal, *layers)
        return scene

    def upload_environment(self, packed) -> None:
        m = packed.meta
        self.environment(packed.env_rgba, m["envWidth"], m["envHeight"], packed.env_table, m["envTableWidth"],
                         m["envTableHeight"])

    def environment(self, rgba_half: np.ndarray, width: int, height: int, table: np.ndarray, tw: int, th: int) -> None:
        rgba = np.ascontiguousarray(rgba_half, dtype=np.uint16)
        tab = np.ascontiguousarray(table, dtype=np.float32)
        _check(load().rm_environment(self.handle, _ptr(rgba), width, height, _ptr(tab), tw, th), "rm_environment")

    def textures(self, scene: Scene, normal_maps: bool, rgba: np.ndarray, width: int, height: int, layers: int):
        data = np.ascontiguousarray(rgba, dtype=np.uint8)
        _check(load().rm_scene_textures(self.handle, scene.handle, int(normal_maps), _ptr(data), width, height, layers),
               "rm_scene_textures")

    def resize(self, width: int, height: int) -> None:
        _check(load().rm_resize(self.handle, width, height), "rm_resize")
        self.width, self.height = width, height

    def frame(self, scene: Scene, params: np.ndarray, emitters: np.ndarray | None = None) -> None:
        p = np.ascontiguousarray(params, dtype=np.uint8)
        e = None if emitters is None else np.ascontiguousarray(emitters, dtype=np.float32)
        _check(load().rm_frame(self.handle, scene.handle, _ptr(p), _ptr(e), 0 if e is None else e.size // 4),
               "rm_frame")

    def read(self, buffer: int = BUFFER_ACCUM) -> np.ndarray:
        out = np.empty((self.height, self.width, 4), dtype=np.float32)
        _check(load().rm_read(self.handle, buffer, _ptr(out), out.size), "rm_read")
        return out

    def display(self) -> np.ndarray:
        """The display pass on the last frame: (H, W, 4) uint8 RGBA."""
        out = np.empty((self.height, self.width, 4), dtype=np.uint8)
        _check(load().rm_display(self.handle, _ptr(out), out.size), "rm_display")
        return out

    def synchronize(self) -> None:
        _check(load().rm_synchronize(self.handle), "rm_synchronize")

    def raycast(self, scene: Scene, rays: np.ndarray, any_hit: bool = False, counters: bool = False):
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        n = r.shape[0]
        hits = np.empty((n, 4), dtype=np.uint32)
        cnt = np.empty((n, 2), dtype=np.uint32) if counters else None
        _check(load().rm_raycast(self.handle, scene.handle, _ptr(r), n, int(any_hit), _ptr(hits), _ptr(cnt)),
               "rm_raycast")
        return (hits, cnt) if counters else hits

    def raycast_static(self, scene: Scene, rays: int, n: int, hits: int, any_hit: bool = False,
                       counters: int | None = None) -> None:
        """rm_raycast_static: rays, hits and counters are device addresses
        (ints); enqueues the kernel on the context's stream without waiting
        and without allocating anything."""
        _check(load().rm_raycast_static(self.handle, scene.handle, C.c_void_p(rays), n, int(any_hit),
                                        C.c_void_p(hits), None if counters is None else C.c_void_p(counters)),
               "rm_raycast_static")

    def tune(self, scene: Scene, params: np.ndarray, frames: int = 2, emitters: np.ndarray | None = None) -> dict:
        """rm_tune: times whole frames under each candidate (schedule,
        persistent blocks, waves, LDS stack depth, block shape, split depth)
        and keeps the fastest; returns {"best": (schedule name, blocks, waves,
        depth, shape, split depth), "candidates": [(schedule name, blocks,
        waves, depth, shape, split depth, ms per frame), ...]}. emitters: as
        for frame(), when the parameters turn light tracing on."""
        res = _TuneResult()
        em = None if emitters is None else np.ascontiguousarray(emitters, dtype=np.float32)
        _check(load().rm_tune(self.handle, scene.handle, _ptr(np.ascontiguousarray(params)), _ptr(em),
                              0 if em is None else em.size // 4, frames, C.byref(res)), "rm_tune")
        names = {v: k for k, v in SCHEDULES.items()}
        cands = [(names[res.schedule[i]], res.persistent_blocks[i], res.waves[i], res.stack_depth[i],
                  "rows" if res.block_shape[i] == BLOCK_ROWS else "squares", res.split_depth[i], res.ms[i])
                 for i in range(res.count)]
        return {"best": cands[res.best][:6], "candidates": cands}

    def set_option(self, option: int, value: int) -> None:
        _check(load().rm_set_option(self.handle, option, value), "rm_set_option")

    def timing(self, enabled: bool) -> None:
        _check(load().rm_timing(self.handle, int(enabled)), "rm_timing")

    def stage_times(self) -> dict:
        n = len(STAGES)
        ms = (C.c_double * n)()
        launches = (C.c_uint32 * n)()
        _check(load().rm_stage_times(self.handle, ms, launches, n), "rm_stage_times")
        return {s: (ms[i], launches[i]) for i, s in enumerate(STAGES)}

    def reset_stage_times(self) -> None:
        _check(load().rm_reset_stage_times(self.handle), "rm_reset_stage_times")

    def device_info(self) -> tuple[str, int]:
        buf = C.create_string_buffer(256)
        cus = C.c_int()
        _check(load().rm_device_info(self.handle, buf, 256, C.byref(cus)), "rm_device_info")
        return buf.value.decode(), cus.value


class _Job(C.Structure):
    _fields_ = [("scene", _P), ("params", _P), ("emitters", _P), ("emitter_vec4s", _U32), ("env_rgba", _P),
                ("env_width", _U32), ("env_height", _U32), ("env_table", _P), ("table_width", _U32),
                ("table_height", _U32), ("target_samples", _U32), ("deadline_ms", C.c_double)]


class _JobStats(C.Structure):
    _fields_ = [("queued_ms", C.c_double), ("first_frame_ms", C.c_double), ("done_ms", C.c_double),
                ("samples", _U32), ("frames", _U32)]


class Server:
    """The job server (rm_server_*): scenes uploaded once on `context`, jobs
    submitted as they arrive, admitted into `slots` concurrent slots within
    `memory_budget` bytes, each job's image and timings returned by wait()."""

    def __init__(self, device: int = 0, slots: int = 8, memory_budget: int = 200 << 30):
        lib = load()
        h = C.c_void_p()
        _check(lib.rm_server_open(device, slots, memory_budget, C.byref(h)), "rm_server_open")
        self.handle = h
        self.context = Context(handle=C.c_void_p(lib.rm_server_context(h)))
        self.sizes = {}

    def submit(self, scene: Scene, params: np.ndarray, target_samples: int, emitters: np.ndarray | None = None,
               packed=None, deadline_ms: float = 0.0) -> int:
        """Queues a job; `packed` (a scenefile.Packed) supplies the environment when it has a map;
        deadline_ms > 0 ends the job (RM_ERR_CANCELLED) if its last frame isn't queued by then."""
        p = np.ascontiguousarray(params, dtype=np.uint8)
        e = None if emitters is None else np.ascontiguousarray(emitters, dtype=np.float32)
        job = _Job(scene=scene.handle, params=_ptr(p), emitters=_ptr(e), emitter_vec4s=0 if e is None else e.size // 4,
                   target_samples=target_samples, deadline_ms=deadline_ms)
        keep = [p, e]
        if packed is not None and packed.meta.get("envTableHeight", 0) > 0:
            rgba = np.ascontiguousarray(packed.env_rgba, dtype=np.uint16)
            table = np.ascontiguousarray(packed.env_table, dtype=np.float32)
            keep += [rgba, table]
            m = packed.meta
            job.env_rgba, job.env_width, job.env_height = _ptr(rgba), m["envWidth"], m["envHeight"]
            job.env_table, job.table_width, job.table_height = _ptr(table), m["envTableWidth"], m["envTableHeight"]
        job_id = C.c_uint64()
        _check(load().rm_server_submit(self.handle, C.byref(job), C.byref(job_id)), "rm_server_submit")
        words = p.view(np.uint32)
        self.sizes[job_id.value] = (int(words[41]), int(words[40]))
        return job_id.value

    def cancel(self, job_id: int) -> None:
        """rm_server_cancel: the job stops (wait() then raises, saying it was cancelled)."""
        _check(load().rm_server_cancel(self.handle, job_id), "rm_server_cancel")

    def set_option(self, option: int, value: int) -> None:
        """An OPTION_* for the jobs submitted from now on."""
        _check(load().rm_server_set_option(self.handle, option, value), "rm_server_set_option")

    def wait(self, job_id: int):
        """(image as (H, W, 4) float32, timing dict) of a job, once it is done."""
        h, w = self.sizes.pop(job_id)
        out = np.empty((h, w, 4), dtype=np.float32)
        st = _JobStats()
        _check(load().rm_server_wait(self.handle, job_id, _ptr(out), out.size, C.byref(st)), "rm_server_wait")
        return out, {"queued_ms": st.queued_ms, "first_frame_ms": st.first_frame_ms, "done_ms": st.done_ms,
                     "samples": st.samples, "frames": st.frames}

    def close(self) -> None:
        if self.handle:
            _check(load().rm_server_close(self.handle), "rm_server_close")
            self.handle = None
            self.context.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
