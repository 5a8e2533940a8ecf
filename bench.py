#!/usr/bin/env python3
"""Raymond on one MI355X: Msamples/s of progressive path tracing at 1080p.

Default workload (spheres200-1920x1080): 200 seeded spheres on a 20 × 20 m
checker floor (50% diffuse, 25% GGX metal, 15% glass, 10% lights), sky
gradient and a 0.5° sun, rendered at 1920 × 1080 with 1 sample per pixel per
frame, 8 bounces, Russian roulette from bounce 3 and NEE+MIS. A step is one
frame exactly as the reference's Renderer.render() issues it with its
defaults (firefly clamp off): the path-tracing pass, light tracing (glass is
present, so caustics are on), splat resolve, photon gather, and the à-trous
denoiser (prepare, 5 passes, compose).

Other workloads (--workload), reported beside it:
  triangles2m-1920x1080  about 2.2M smooth triangles (fBm terrain, tori,
                         lathes, Bézier tubes) in GGX plastic, a fog box and a
                         2048 × 1024 environment map, 12 bounces, 1 spp/frame
                         (volumes turn light tracing off, as in the reference);
  caustics-1920x1080     a 512 × 512-vertex water sheet, a dispersive glass
                         ball, a spherical lamp and the sun, light tracing and
                         photon gather every frame;
  cornell-raycast-256    the 256 × 256 pinhole primary rays of the Cornell box,
                         closest hit, in Mrays/s, checked bit for bit against
                         the oracle;
  serving-poisson        jobs arriving as a Poisson process (scene from a pool
                         of 9 to 2.2 M primitives, 256² to 1080p, 4-12
                         bounces, a target spp), served by the job server with
                         concurrent slots and, on the same trace, one job at a
                         time: throughput and p50/p99 latencies of both.

Prints one JSON line; `roofline` is for the dominant kernel (the path-tracing
pass, or the raycast kernel), `cpu_baseline` times the CPU oracle on the
box's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# Non-packed VALU issue peak of MI355X: 256 CUs × 4 SIMDs × 16 lanes × 2.4 GHz
# = 39.3 T lane-operations/s (the datasheet's 157.3 TFLOP/s FP32 counts an
# FMA as two flops and packed-FP32 issue, which the kernels are compiled without).
VALU_PEAK_GLANE_OPS = 256 * 64 * 2.4
SIMDS = 1024
METRIC = "Msamples/s at 1080p 8-bounce; served p50/p99 latency vs plain, 1 MI355X"

WORKLOADS = {
    "spheres200-1920x1080": {
        "config": "spheres", "bounces": 8, "stride": 31,
        "opts": {"scene": {"seed": 2, "count": 200, "width": 1920, "height": 1080}, "maxBounces": 8,
                 "samplesPerFrame": 1, "fireflyClamp": 0.0, "denoise": True, "caustics": True},
        "data": "synthetic: seeded 200-sphere scene generated and packed by the JS host modules",
        # The path pass's counter summaries, by schedule.
        "traffic": {"split": ROOT / "profiles" / "k_split_spheres200_1080p_pmc.md",
                    "per-pixel": ROOT / "profiles" / "k_main_spheres200_1080p_pmc.md"},
    },
    "triangles2m-1920x1080": {
        "config": "triangles", "bounces": 12, "stride": 509,
        "opts": {"scene": {"seed": 3, "grid": 1024, "width": 1920, "height": 1080}, "maxBounces": 12,
                 "samplesPerFrame": 1, "fireflyClamp": 0.0, "denoise": True, "caustics": True},
        "data": "synthetic: seeded terrain/tori/lathes/tubes scene with fog and a generated environment map, "
                "packed by the JS host modules",
        "traffic": None,
    },
    "caustics-1920x1080": {
        "config": "caustics", "bounces": 8, "stride": 127,
        "opts": {"scene": {"seed": 4, "res": 512, "width": 1920, "height": 1080}, "maxBounces": 8,
                 "samplesPerFrame": 1, "fireflyClamp": 0.0, "denoise": True, "caustics": True},
        "data": "synthetic: seeded water sheet, dispersive glass ball, lamp and sun, packed by the JS host modules",
        "traffic": None,
    },
}
# The caustics workload as described for it: 8M light paths a frame (each
# depositing at most one photon, so 8M photon slots; the reference traces
# min(W·H·spp, 524,288)) and a fixed 1 cm gather radius (the reference's
# shrinks as frame^(-1/6)).
WORKLOADS["caustics-8m-photons-1920x1080"] = dict(
    WORKLOADS["caustics-1920x1080"], light_paths=8 << 20, photon_radius=0.01,
    data="synthetic: seeded water sheet, dispersive glass ball, lamp and sun, packed by the JS host modules; "
         "8M light paths and photon slots a frame, 1 cm gather radius")
RAYCAST = "cornell-raycast-256"
SERVING = "serving-poisson"

# The served trace's scene pool, smallest to largest (9 to 2.2 M primitives).
SERVING_POOL = [
    ("cornell", {"width": 256, "height": 256}),
    ("spheres", {"scene": {"width": 256, "height": 256}, "width": 256, "height": 256}),
    ("mixed", {"width": 256, "height": 256}),
    ("caustics", {"scene": {"res": 128, "width": 256, "height": 256}, "width": 256, "height": 256}),
    ("triangles", {"scene": {"grid": 256, "width": 256, "height": 256}, "width": 256, "height": 256}),
    ("triangles", {"scene": {"grid": 1024, "width": 256, "height": 256}, "width": 256, "height": 256}),
]
SERVING_SIZES = [(256, 256), (512, 512), (1280, 720), (1920, 1080)]


def summary_figure(path, key):
    """A per-launch figure (`key: value`) from a committed PMC summary."""
    if path is None or not Path(path).exists():
        return None
    m = re.search(rf"{key}:\s*([0-9.eE+]+)", Path(path).read_text())
    return float(m.group(1)) if m else None


def traffic_from_summary(path):
    """HBM bytes per launch of the trace kernel, from a committed PMC summary."""
    return summary_figure(path, "hbm_bytes_per_launch")


def cpu_model() -> str:
    """The host CPU's model name (lscpu), for the CPU baseline."""
    import subprocess
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        m = re.search(r"Model name:\s*(.+)", out)
        if m:
            return m.group(1).strip()
    except (OSError, subprocess.SubprocessError):
        pass
    for line in Path("/proc/cpuinfo").read_text().splitlines():
        if line.startswith("model name"):
            return line.split(":", 1)[1].strip()
    return "unknown"


def roofline(kernel_s: float, algorithmic_bytes: float, summary) -> dict:
    """The path pass against both rooflines that could bound it.

    VALU: lane-operations per launch (VALU wave instructions × mean active
    lanes, from SQ_INSTS_VALU, SQ_THREAD_CYCLES_VALU and SQ_ACTIVE_INST_VALU
    in the committed PMC summary) per second of the launch, against the
    non-packed issue peak. HBM: the measured FETCH_SIZE + WRITE_SIZE bytes of
    a launch per second against 8 TB/s. `bound` names the larger fraction.
    The algorithmic bytes (node, primitive, material and per-pixel records)
    mostly hit L2; they are reported as the working-set rate, not as HBM."""
    lane_ops = summary_figure(summary, "valu_lane_ops_per_launch")
    issue = summary_figure(summary, "valu_insts_per_launch")
    traffic = traffic_from_summary(summary)
    hbm = None if traffic is None else traffic / kernel_s / 1e9
    valu = None if lane_ops is None else lane_ops / kernel_s / 1e9
    fracs = {"valu": None if valu is None else valu / VALU_PEAK_GLANE_OPS,
             "hbm": None if hbm is None else hbm / HBM_PEAK_GBS}
    if fracs["valu"] is not None and (fracs["hbm"] is None or fracs["valu"] >= fracs["hbm"]):
        line = {"bound": "valu", "achieved": round(valu, 1), "peak": round(VALU_PEAK_GLANE_OPS, 1),
                "unit": "Glane-op/s", "frac": round(fracs["valu"], 4)}
    else:
        alg = algorithmic_bytes / kernel_s / 1e9
        line = {"bound": "hbm", "achieved": round(hbm if hbm is not None else alg, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round((hbm if hbm is not None else alg) / HBM_PEAK_GBS, 4),
                "basis": "measured HBM traffic" if hbm is not None else
                         "algorithmic bytes (no counter summary for this workload; mostly L2 hits)"}
    line["traffic"] = traffic
    line["valu"] = None if valu is None else {
        "achieved_glane_ops": round(valu, 1), "frac": round(fracs["valu"], 4),
        "lanes_active_of_64": round(lane_ops / issue, 2) if issue else None,
        "issue_busy": round(issue * 4 / (SIMDS * 2.4e9 * kernel_s), 4) if issue else None}
    line["hbm"] = None if hbm is None else {"achieved_gbs": round(hbm, 1), "frac": round(fracs["hbm"], 4)}
    line["l2_working_set"] = {"algorithmic_bytes_per_launch": round(algorithmic_bytes),
                              "gbs": round(algorithmic_bytes / kernel_s / 1e9, 1),
                              "note": "node, primitive, material and per-pixel record bytes the traversal and "
                                      "shading read; the scene stays in L2, so this is not HBM traffic"}
    return line


def algorithmic_bytes(p, params, stride: int) -> dict:
    """Per-sample algorithmic bytes of the path-tracing pass, from the oracle's
    segment counts on a pixel subsample of the same frame:
      rays:     24 B in + 16 B hit out (closest) or 4 B (shadow)
      traversal 32 B per node record tested, 64 B per primitive record tested
      shading:  64 B primitive + 96 B material per surface hit, +64 B attributes for smooth triangles
      per pixel per frame: 3 × 16 B read + 3 × 16 B written (accumulation and guides)."""
    from oracle.pyoracle import Oracle
    st = Oracle(p).path_stats(params, stride)
    n = st["samples"]
    per_sample = (st["closest"] * (24 + 16) + st["shadow"] * (24 + 4) + 32 * st["nodes"] + 64 * st["prims"]
                  + st["surfaces"] * (64 + 96) + 64 * st["smooth"]) / n
    return {"per_sample": per_sample, "rays_per_sample": (st["closest"] + st["shadow"]) / n, "stats": st}


def cpu_threads() -> int:
    return int(os.environ.get("RAYMOND_CPU_THREADS", os.cpu_count() or 1))


def cpu_paths(p, params_for_frame, threads: int, budget_s: float, rows: int) -> tuple:
    """Times the oracle's main() (camera paths) on bands of `rows` rows of
    successive frames until about budget_s seconds: (samples, seconds, bands)."""
    from oracle.pyoracle import Oracle
    o = Oracle(p)
    w, h = p.width, p.height
    acc, ga, gn = (np.zeros((h, w, 4), np.float32) for _ in range(3))
    bands = 0
    t0 = time.perf_counter()
    while True:
        frame, band = divmod(bands, max(h // rows, 1))
        y0 = band * rows
        o.main(params_for_frame(frame), acc, ga, gn, rows=(y0, min(h, y0 + rows)), threads=threads)
        bands += 1
        dt = time.perf_counter() - t0
        if dt >= budget_s:
            return bands * rows * w, dt, bands


def cpu_baseline(p, params_for_frame, budget_s: float = 10.0) -> dict:
    """The oracle's main() (camera paths only, the same arithmetic and RNG
    streams) on the host's cores, then on one: `value` is the all-cores
    rate, `single_thread` the one-core rate, each over bands of whole rows
    of the same 1080p frames for about budget_s seconds."""
    threads = cpu_threads()
    w, h = p.width, p.height
    samples, dt, bands = cpu_paths(p, params_for_frame, threads, budget_s, rows=h // 4)
    one, dt1, bands1 = cpu_paths(p, params_for_frame, 1, budget_s / 2, rows=8)
    return {"value": round(samples / dt / 1e6, 4), "unit": "Msamples/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"oracle main() camera paths over {bands} quarter-frames of {w}x{h} at 1 spp ({samples} "
                      f"samples, {dt:.1f} s, {threads} threads); light tracing, photon gather and denoise not "
                      f"included",
            "single_thread": {"value": round(one / dt1 / 1e6, 4), "unit": "Msamples/s", "cores": 1,
                              "sample": f"{bands1} bands of 8 rows ({one} samples, {dt1:.1f} s)"}}


def bench_frames(args, name: str) -> dict:
    from raymond_amd import capi, scenefile
    wl = WORKLOADS[name]
    p = scenefile.pack(wl["config"], wl["opts"])
    w, h = p.width, p.height
    spp = scenefile.get_param(p.params, "samples_per_frame")
    r0 = scenefile.get_param(p.params, "photon_radius")
    caustics = scenefile.get_param(p.params, "caustics")

    def frame_params(k: int) -> np.ndarray:
        r = wl.get("photon_radius") or r0 * (k + 1) ** (-1.0 / 6.0)
        return p.with_params(sample_offset=k * spp, total_samples=(k + 1) * spp, photon_radius=r,
                             photon_cell=2 * r)

    ctx = capi.Context(0)
    device, cus = ctx.device_info()
    tuned = args.schedule == "tuned"
    if not tuned:
        ctx.set_option(capi.OPTION_SCHEDULE, capi.SCHEDULES[args.schedule])
        ctx.set_option(capi.OPTION_PERSISTENT_BLOCKS, args.persistent_blocks)
        ctx.set_option(capi.OPTION_WAVES, args.waves)
        ctx.set_option(capi.OPTION_SPLIT_DEPTH, args.split_depth)
    ctx.set_option(capi.OPTION_PRECISION, capi.PRECISION_QUANTIZED_BVH if args.quantized_bvh else 0)
    ctx.set_option(capi.OPTION_LIGHT_PATHS, wl.get("light_paths", 0))
    scene = ctx.upload_scene(p)
    ctx.environment(p.env_rgba, p.meta["envWidth"], p.meta["envHeight"], p.env_table, p.meta["envTableWidth"],
                    p.meta["envTableHeight"])
    ctx.resize(w, h)
    emitters = p.emitters if caustics else None
    # Tuned with the first timed frame's parameters: its photon radius, not the first frame's (which gathers
    # over about twice the area), sets the side stream's load the path pass shares the GPU with.
    tuning = (ctx.tune(scene, frame_params(args.warmup), frames=args.tune_frames, emitters=emitters)
              if tuned else None)

    k = 0
    for _ in range(args.warmup):
        ctx.frame(scene, frame_params(k), emitters)
        k += 1
    ctx.synchronize()
    ctx.timing(True)
    ctx.reset_stage_times()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.frame(scene, frame_params(k), emitters)
        k += 1
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    stages = ctx.stage_times()
    ctx.timing(False)
    scene.release()
    ctx.close()

    samples = w * h * spp * args.steps
    value = samples / elapsed / 1e6
    schedule = tuning["best"][0] if tuned else args.schedule
    # The committed counter summaries were taken at four waves, split depth 2, on the FP32 tree.
    waves = tuning["best"][2] if tuned else args.waves
    split_depth = (tuning["best"][5] if tuned else args.split_depth) or 2
    as_profiled = waves == 4 and split_depth == 2 and not args.quantized_bvh
    summary = (wl["traffic"] or {}).get(schedule) if as_profiled else None
    trace_ms, trace_n = stages["trace"]
    trace_avg_s = trace_ms / max(trace_n, 1) / 1e3

    alg = algorithmic_bytes(p, frame_params(0), args.stride or wl["stride"])
    bytes_per_launch = alg["per_sample"] * w * h * spp + 96 * w * h
    bounces = wl["bounces"]
    line = {
        "metric": METRIC if name == "spheres200-1920x1080" else f"Msamples/s at 1080p {bounces}-bounce, 1 MI355X",
        "value": round(value, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": wl["data"],
        "config": {"workload": name, "width": w, "height": h, "spp_per_frame": spp, "max_bounces": bounces,
                   "rr_from_bounce": 3, "nee_mis": True, "light_tracing": bool(caustics),
                   "photon_gather": bool(caustics), "denoise_passes": 5, "frames_timed": args.steps,
                   "primitives": p.primitive_count, "device": device, "compute_units": cus,
                   "light_paths_per_frame": wl.get("light_paths") or min(w * h * spp, 524288) if caustics else 0,
                   "photon_radius": wl.get("photon_radius", "shrinking as frame^(-1/6) from a 3-pixel footprint"),
                   "path": "one job, " + ("8-bit quantized BVH" if args.quantized_bvh else "FP32 BVH"),
                   "schedule": args.schedule, "waves_per_simd": waves,
                   "split_depth": split_depth if schedule == "split" else None,
                   "tuning": tuning},
        "roofline": dict(roofline(trace_avg_s, bytes_per_launch, summary),
                         kernel={"compact": "k_main_compact", "sorted": "k_main_compact (sorting)",
                                 "per-pixel": "k_main", "persistent": "k_main_persistent",
                                 "wavefront": "the wavefront stages (k_wf_*)",
                                 "split": "k_split_head + k_split_tail"}[schedule],
                         launch_ms=round(trace_avg_s * 1e3, 4),
                         bytes_per_sample=round(alg["per_sample"], 1),
                         rays_per_sample=round(alg["rays_per_sample"], 3)),
        "stages_ms_per_frame": {s: round(ms / max(n, 1), 4) for s, (ms, n) in stages.items() if n},
    }
    if name == "spheres200-1920x1080" and not args.no_serving:
        line["served"] = served_vs_plain(args)
    line["cpu_baseline"] = None if args.no_cpu else cpu_baseline(p, frame_params)
    return line


def bench_raycast(args) -> dict:
    """Closest-hit primary rays of the Cornell box, bit-exact against the oracle."""
    from oracle.pyoracle import Oracle
    from raymond_amd import capi, scenefile
    p = scenefile.pack("cornell", {"width": 256, "height": 256})
    o = Oracle(p)
    rays = o.camera_rays()
    n = rays.shape[0]
    ctx = capi.Context(0)
    ctx.set_option(capi.OPTION_PRECISION, capi.PRECISION_QUANTIZED_BVH if args.quantized_bvh else 0)
    device, cus = ctx.device_info()
    scene = ctx.upload_scene(p)
    hits, counters = ctx.raycast(scene, rays, counters=True)
    exact = bool(np.array_equal(hits, o.trace(rays)))
    for _ in range(args.warmup):
        ctx.raycast(scene, rays)
    ctx.timing(True)
    ctx.reset_stage_times()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.raycast(scene, rays)
    elapsed = time.perf_counter() - t0
    ms, launches = ctx.stage_times()["raycast"]
    ctx.timing(False)
    scene.release()
    ctx.close()
    kernel_s = ms / max(launches, 1) / 1e3
    # 24 B ray in, 16 B hit out, 32 B per node record and 64 B per primitive record tested.
    bytes_per_launch = n * (24 + 16) + 32 * int(counters[:, 0].sum()) + 64 * int(counters[:, 1].sum())
    achieved = bytes_per_launch / kernel_s / 1e9
    threads = cpu_threads()
    t0 = time.perf_counter()
    reps = 0
    while time.perf_counter() - t0 < 10.0:
        o.trace(rays, threads=threads)
        reps += 1
    cpu_s = (time.perf_counter() - t0) / reps
    return {
        "metric": "Mrays/s, 256x256 Cornell primary rays, closest hit, 1 MI355X",
        "value": round(n / kernel_s / 1e6, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: Cornell box (5 walls, lamp, two yawed boxes, glass ball) packed by the JS host modules",
        "config": {"workload": RAYCAST, "rays": n, "bit_exact_vs_oracle": exact, "device": device,
                   "compute_units": cus,
                   "note": "value is kernel time only; ms_per_step includes the host-to-device ray upload "
                           "and hit readback of each call"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "k_raycast",
                     "launch_ms": round(kernel_s * 1e3, 4), "algorithmic_bytes_per_launch": bytes_per_launch},
        "cpu_baseline": None if args.no_cpu else {
            "value": round(n / cpu_s / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"oracle trace() over the same {n} rays, {reps} repetitions in {reps * cpu_s:.1f} s"},
    }


def serving_trace(rate: float, duration: float, spp_lo: int, spp_hi: int, seed: int = 5) -> list:
    """Poisson arrivals at `rate` jobs/s over `duration` s: (arrival s, scene, size, target spp, bounces)."""
    rng = np.random.default_rng(seed)
    jobs, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return jobs
        jobs.append((t, int(rng.integers(len(SERVING_POOL))), SERVING_SIZES[int(rng.integers(len(SERVING_SIZES)))],
                     int(rng.integers(spp_lo, spp_hi + 1)), int(rng.integers(4, 13))))


PRECISION_FLAGS = {"quantized-bvh": 1, "fp16-guides": 2, "compensated": 4}


def precision_flags(names: str) -> int:
    flags = 0
    for n in filter(None, names.split(",")):
        flags |= PRECISION_FLAGS[n]
    return flags


def run_trace(packs, trace, slots: int, budget: int, precision: int = 0, autotune: int = 0, batch: bool = False,
              schedule: int | None = None, tune_mask: int = 0) -> dict:
    """Submits the trace's jobs at their arrival times to a server with `slots`
    slots, RM_OPTION_PRECISION flags `precision` and the schedule given (None:
    the server's default, the one the plain path's context starts with), and
    with `autotune` frames a tuner candidate over the candidates `tune_mask`
    names (0: all); returns throughput and latency percentiles."""
    from raymond_amd import capi, scenefile
    with capi.Server(slots=slots, memory_budget=budget) as server:
        server.set_option(capi.OPTION_PRECISION, precision)
        server.set_option(capi.OPTION_AUTOTUNE, autotune)
        server.set_option(capi.OPTION_TUNE_CANDIDATES, tune_mask)
        server.set_option(capi.OPTION_BATCH, int(batch))
        if schedule is not None:
            server.set_option(capi.OPTION_SCHEDULE, schedule)
        scenes = [server.context.upload_scene(p) for p in packs]
        jobs = []
        for arrival, k, (w, h), spp, bounces in trace:
            p = packs[k]
            params = scenefile.resized_params(p.params, w, h, bounces)
            caustics = bool(scenefile.get_param(params, "caustics"))
            jobs.append((arrival, scenes[k], params, spp, p.emitters if caustics else None, p, w * h * spp))
        t0 = time.perf_counter()
        ids = []
        for arrival, scene, params, spp, emitters, p, _ in jobs:
            delay = t0 + arrival - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            ids.append(server.submit(scene, params, spp, emitters, packed=p))
        stats = [server.wait(i)[1] for i in ids]
        elapsed = time.perf_counter() - t0
    samples = sum(j[-1] for j in jobs)
    first = np.array([s["first_frame_ms"] for s in stats])
    done = np.array([s["done_ms"] for s in stats])
    return {"Msamples_per_s": round(samples / elapsed / 1e6, 3), "wall_s": round(elapsed, 3), "jobs": len(jobs),
            "samples": samples, "slots": slots, "precision": precision, "batch": batch,
            "first_frame_ms_p50": round(float(np.percentile(first, 50)), 2),
            "first_frame_ms_p99": round(float(np.percentile(first, 99)), 2),
            "done_ms_p50": round(float(np.percentile(done, 50)), 2),
            "done_ms_p99": round(float(np.percentile(done, 99)), 2)}


def run_plain(packs, trace) -> dict:
    """The same trace on the plain path: one context, jobs one at a time in
    arrival order, every frame rendered and denoised as the Renderer renders
    it, the first frame waited for (the image a viewer first sees), the last
    read back. Latencies count from each job's arrival."""
    from raymond_amd import capi, scenefile
    first, done, samples = [], [], 0
    with capi.Context(0) as ctx:
        scenes = [ctx.upload_scene(p) for p in packs]
        ctx.synchronize()
        t0 = time.perf_counter()
        for arrival, k, (w, h), spp, bounces in trace:
            delay = t0 + arrival - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            p = packs[k]
            params = scenefile.resized_params(p.params, w, h, bounces)
            caustics = bool(scenefile.get_param(params, "caustics"))
            r0 = scenefile.get_param(params, "photon_radius")
            ctx.upload_environment(p)
            ctx.resize(w, h)
            for f in range(spp):
                r = r0 * (f + 1) ** (-1.0 / 6.0)
                fp = scenefile.set_params(params.copy(), sample_offset=f, total_samples=f + 1, photon_radius=r,
                                          photon_cell=2 * r)
                ctx.frame(scenes[k], fp, p.emitters if caustics else None)
                if f == 0:
                    ctx.synchronize()
                    first.append((time.perf_counter() - t0 - arrival) * 1e3)
            ctx.read(capi.BUFFER_DENOISED)
            done.append((time.perf_counter() - t0 - arrival) * 1e3)
            samples += w * h * spp
        elapsed = time.perf_counter() - t0
    first, done = np.array(first), np.array(done)
    return {"Msamples_per_s": round(samples / elapsed / 1e6, 3), "wall_s": round(elapsed, 3), "jobs": len(trace),
            "samples": samples, "slots": "plain",
            "first_frame_ms_p50": round(float(np.percentile(first, 50)), 2),
            "first_frame_ms_p99": round(float(np.percentile(first, 99)), 2),
            "done_ms_p50": round(float(np.percentile(done, 50)), 2),
            "done_ms_p99": round(float(np.percentile(done, 99)), 2)}


def serving_packs():
    from raymond_amd import scenefile
    return [scenefile.pack(c, dict(o, maxBounces=12, samplesPerFrame=1, fireflyClamp=0.0, denoise=True,
                                   caustics=True)) for c, o in SERVING_POOL]


def served_vs_plain(args) -> dict:
    """The served half of the metric on a short trace: Poisson arrivals at
    args.rate jobs/s for args.serve_seconds s over the scene pool, served
    (8 slots, FP32), served in the form the served workload describes
    (batched path passes, quantized BVH, FP16 guides, compensated sums), and
    the plain path one job at a time; Msamples/s and p50/p99 time to the
    first denoised frame and to the target samples of each."""
    packs = serving_packs()
    trace = serving_trace(args.rate, args.serve_seconds, args.spp_min, args.spp_max)
    budget = 200 << 30
    served = run_trace(packs, trace, args.slots, budget)
    described = run_trace(packs, trace, args.slots, budget, precision_flags(args.served_precision), batch=True)
    plain = run_plain(packs, trace)
    # At the described rate the scaled-down trace is arrival-bound (both
    # paths keep up with the arrivals); a rate that saturates both shows the
    # capacity of each.
    busy = serving_trace(args.capacity_rate, args.serve_seconds, args.spp_min, args.spp_max)
    busy_served = run_trace(packs, busy, args.slots, budget)
    busy_plain = run_plain(packs, busy)
    keep = ("Msamples_per_s", "first_frame_ms_p50", "first_frame_ms_p99", "done_ms_p50", "done_ms_p99", "wall_s")
    return {"trace": {"rate_jobs_per_s": args.rate, "arrivals_s": args.serve_seconds, "jobs": len(trace),
                      "spp_target": [args.spp_min, args.spp_max], "bounces": [4, 12],
                      "sizes": [f"{w}x{h}" for w, h in SERVING_SIZES],
                      "scene_primitives": [p.primitive_count for p in packs],
                      "scaled_down": f"{args.serve_seconds:g} s of arrivals and spp U{{{args.spp_min}..{args.spp_max}}} "
                                     "rather than 600 s and U{16..1024}, so the run fits the bench's minute"},
            "served": {k: served[k] for k in keep},
            "served_batched_precision_variants": {k: described[k] for k in keep},
            "plain": {k: plain[k] for k in keep},
            "speedup_vs_plain": round(served["Msamples_per_s"] / max(plain["Msamples_per_s"], 1e-9), 3),
            "capacity": {"rate_jobs_per_s": args.capacity_rate, "jobs": len(busy),
                         "served": {k: busy_served[k] for k in keep}, "plain": {k: busy_plain[k] for k in keep},
                         "speedup_vs_plain": round(busy_served["Msamples_per_s"] /
                                                   max(busy_plain["Msamples_per_s"], 1e-9), 3)}}


SERVING_VARIANTS = ("served", "served_batched", "served_batched_wavefront", "served_tuned", "served_tuned_split",
                    "served_batched_with_precision_variants", "served_one_slot", "plain")


def bench_serving(args) -> dict:
    """A Poisson trace of jobs over the scene pool, served with concurrent
    slots and, on the same trace, one job at a time (the plain path)."""
    from raymond_amd import capi
    packs = serving_packs()
    trace = serving_trace(args.rate, args.duration, args.spp_min, args.spp_max)
    budget = 200 << 30
    flags = precision_flags(args.served_precision)
    wanted = set(args.variants.split(","))
    unknown = wanted - set(SERVING_VARIANTS)
    if unknown or "served" not in wanted:
        raise SystemExit(f"--variants takes 'served' and any of {', '.join(SERVING_VARIANTS)}; got {args.variants}")
    runs = {
        "served": lambda: run_trace(packs, trace, args.slots, budget),
        "served_batched": lambda: run_trace(packs, trace, args.slots, budget, batch=True),
        "served_batched_wavefront": lambda: run_trace(packs, trace, args.slots, budget, batch=True,
                                                      schedule=capi.SCHEDULE_WAVEFRONT),
        "served_tuned": lambda: run_trace(packs, trace, args.slots, budget, autotune=1),
        # The tuner's three split candidates only (four and five waves after two bounces, five after one).
        "served_tuned_split": lambda: run_trace(packs, trace, args.slots, budget, autotune=1, tune_mask=0b111),
        "served_batched_with_precision_variants": lambda: run_trace(packs, trace, args.slots, budget, flags,
                                                                    batch=True),
        "served_one_slot": lambda: run_trace(packs, trace, 1, budget),
        "plain": lambda: run_plain(packs, trace),
    }
    results = {}
    for k in SERVING_VARIANTS:
        if k in wanted:
            results[k] = runs[k]()
            print(f"serving: {k} done, {results[k]['Msamples_per_s']} Msamples/s in {results[k]['wall_s']} s",
                  file=sys.stderr, flush=True)
    served, plain = results["served"], results.get("plain")
    with capi.Context(0) as ctx:
        device, cus = ctx.device_info()
    return {
        "metric": "served Msamples/s on a Poisson job trace, with p50/p99 time to first denoised frame and to "
                  "target samples, vs the plain one-job-at-a-time path, 1 MI355X",
        "value": served["Msamples_per_s"],
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": len(trace),
        "warmup": 0,
        "ms_per_step": round(served["wall_s"] * 1e3 / max(len(trace), 1), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: Poisson arrivals over a pool of seeded scenes packed by the JS host modules",
        "config": {"workload": SERVING, "rate_jobs_per_s": args.rate, "duration_s": args.duration,
                   "spp_target": [args.spp_min, args.spp_max], "bounces": [4, 12],
                   "sizes": [f"{w}x{h}" for w, h in SERVING_SIZES],
                   "scene_primitives": [p.primitive_count for p in packs], "slots": args.slots,
                   "memory_budget_bytes": budget, "device": device, "compute_units": cus,
                   "served_precision": args.served_precision,
                   "path": "served: concurrent slots, FP32, a job's first and last frames denoised (also reported: "
                           "the running jobs' path passes batched into one launch; batched on the wavefront schedule, "
                           "each stage one launch for all running jobs, replayed from a hipGraph; autotuned per scene "
                           "and size; and batched with the precision variants named); plain: one context, one job at "
                           "a time, FP32, every frame denoised"},
        **results,
        "speedup_vs_plain": round(served["Msamples_per_s"] / max(plain["Msamples_per_s"], 1e-9), 3) if plain else None,
        "cpu_baseline": None,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--stride", type=int, default=0, help="pixel subsample stride for algorithmic bytes")
    ap.add_argument("--workload", choices=[*WORKLOADS, RAYCAST, SERVING], default="spheres200-1920x1080")
    ap.add_argument("--rate", type=float, default=20.0, help="serving: job arrivals per second")
    ap.add_argument("--duration", type=float, default=15.0, help="serving: seconds of arrivals")
    ap.add_argument("--spp-min", type=int, default=4, help="serving: smallest target spp")
    ap.add_argument("--spp-max", type=int, default=64, help="serving: largest target spp")
    ap.add_argument("--slots", type=int, default=8, help="serving: concurrent job slots")
    ap.add_argument("--serve-seconds", type=float, default=4.0,
                    help="default workload: seconds of arrivals of the short served-vs-plain trace")
    ap.add_argument("--no-serving", action="store_true", help="default workload: skip the served-vs-plain trace")
    ap.add_argument("--capacity-rate", type=float, default=60.0,
                    help="default workload: arrival rate (jobs/s) of the saturating served-vs-plain trace")
    ap.add_argument("--variants", default=",".join(SERVING_VARIANTS),
                    help="serving: the runs to make, comma-separated (served is always one of them)")
    ap.add_argument("--served-precision", default="quantized-bvh,fp16-guides,compensated",
                    help="serving: precision variants of the batched variant run, comma-separated "
                         "(quantized-bvh, fp16-guides, compensated)")
    ap.add_argument("--schedule", choices=["per-pixel", "compact", "sorted", "persistent", "wavefront", "split",
                                           "tuned"],
                    default="split",
                    help="path-pass schedule: one thread per pixel, live paths packed between bounces, "
                         "lanes regenerating paths from a tile queue, or a kernel a pixel for the first bounces "
                         "and persistent waves over the paths still alive")
    ap.add_argument("--tune-frames", type=int, default=16,
                    help="--schedule tuned: synchronised frames a candidate (each candidate is timed twice)")
    ap.add_argument("--split-depth", type=int, default=0,
                    help="bounces the split schedule's first kernel runs (RM_OPTION_SPLIT_DEPTH; 0: 2)")
    ap.add_argument("--waves", type=int, choices=[4, 5], default=4,
                    help="waves per SIMD of the one-thread-per-pixel path pass (RM_OPTION_WAVES)")
    ap.add_argument("--quantized-bvh", action="store_true",
                    help="traverse the 8-bit child-quantized tree (same hits as the FP32 tree)")
    ap.add_argument("--persistent-blocks", type=int, default=0, help="blocks for the persistent pass (0: fill the GPU)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the library has no multi-GPU path (asked for {args.gpus}).", file=sys.stderr)
        return 2
    if args.workload == RAYCAST:
        line = bench_raycast(args)
    elif args.workload == SERVING:
        line = bench_serving(args)
    else:
        line = bench_frames(args, args.workload)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
