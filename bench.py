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
                         the oracle.

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
METRIC = "Msamples/s at 1080p 8-bounce; served p50/p99 latency vs plain, 1 MI355X"

WORKLOADS = {
    "spheres200-1920x1080": {
        "config": "spheres", "bounces": 8, "stride": 31,
        "opts": {"scene": {"seed": 2, "count": 200, "width": 1920, "height": 1080}, "maxBounces": 8,
                 "samplesPerFrame": 1, "fireflyClamp": 0.0, "denoise": True, "caustics": True},
        "data": "synthetic: seeded 200-sphere scene generated and packed by the JS host modules",
        "traffic": ROOT / "profiles" / "k_main_spheres200_1080p_pmc.md",
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
RAYCAST = "cornell-raycast-256"


def traffic_from_summary(path):
    """HBM bytes per launch of the trace kernel, from a committed PMC summary."""
    if path is None or not Path(path).exists():
        return None
    m = re.search(r"hbm_bytes_per_launch:\s*([0-9.eE+]+)", Path(path).read_text())
    return float(m.group(1)) if m else None


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


def cpu_baseline(p, params_for_frame, budget_s: float = 12.0) -> dict:
    """The oracle's main() (camera paths only) over whole frames of the same
    workload, as many as fit in about budget_s seconds of wall time."""
    from oracle.pyoracle import Oracle
    o = Oracle(p)
    threads = cpu_threads()
    w, h = p.width, p.height
    shape = (h, w, 4)
    acc, ga, gn = (np.zeros(shape, np.float32) for _ in range(3))
    frames = 0
    t0 = time.perf_counter()
    while True:
        o.main(params_for_frame(frames), acc, ga, gn, threads=threads)
        frames += 1
        dt = time.perf_counter() - t0
        if dt >= budget_s or dt / frames * (frames + 1) > 2.5 * budget_s:
            break
    samples = frames * w * h
    return {"value": round(samples / dt / 1e6, 4), "unit": "Msamples/s", "cores": threads, "kind": "port",
            "sample": f"oracle main() camera paths: {frames} whole {w}x{h} frames at 1 spp ({samples} samples, "
                      f"{dt:.1f} s); light tracing, photon gather and denoise not included"}


def bench_frames(args, name: str) -> dict:
    from raymond_amd import capi, scenefile
    wl = WORKLOADS[name]
    p = scenefile.pack(wl["config"], wl["opts"])
    w, h = p.width, p.height
    spp = scenefile.get_param(p.params, "samples_per_frame")
    r0 = scenefile.get_param(p.params, "photon_radius")
    caustics = scenefile.get_param(p.params, "caustics")

    def frame_params(k: int) -> np.ndarray:
        r = r0 * (k + 1) ** (-1.0 / 6.0)
        return p.with_params(sample_offset=k * spp, total_samples=(k + 1) * spp, photon_radius=r,
                             photon_cell=2 * r)

    ctx = capi.Context(0)
    device, cus = ctx.device_info()
    ctx.set_option(capi.OPTION_SCHEDULE, capi.SCHEDULES[args.schedule])
    ctx.set_option(capi.OPTION_PERSISTENT_BLOCKS, args.persistent_blocks)
    scene = ctx.upload_scene(p)
    ctx.environment(p.env_rgba, p.meta["envWidth"], p.meta["envHeight"], p.env_table, p.meta["envTableWidth"],
                    p.meta["envTableHeight"])
    ctx.resize(w, h)
    emitters = p.emitters if caustics else None

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
    trace_ms, trace_n = stages["trace"]
    trace_avg_s = trace_ms / max(trace_n, 1) / 1e3

    alg = algorithmic_bytes(p, frame_params(0), args.stride or wl["stride"])
    bytes_per_launch = alg["per_sample"] * w * h * spp + 96 * w * h
    achieved = bytes_per_launch / trace_avg_s / 1e9
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
                   "path": "plain (one job, megakernel, FP32 BVH)", "schedule": args.schedule},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_summary(wl["traffic"]),
                     "kernel": {"compact": "k_main_compact", "per-pixel": "k_main",
                                "persistent": "k_main_persistent"}[args.schedule],
                     "launch_ms": round(trace_avg_s * 1e3, 4),
                     "algorithmic_bytes_per_launch": round(bytes_per_launch),
                     "bytes_per_sample": round(alg["per_sample"], 1),
                     "rays_per_sample": round(alg["rays_per_sample"], 3)},
        "stages_ms_per_frame": {s: round(ms / max(n, 1), 4) for s, (ms, n) in stages.items() if n},
    }
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


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--stride", type=int, default=0, help="pixel subsample stride for algorithmic bytes")
    ap.add_argument("--workload", choices=[*WORKLOADS, RAYCAST], default="spheres200-1920x1080")
    ap.add_argument("--schedule", choices=["per-pixel", "compact", "persistent"], default="per-pixel",
                    help="path-pass schedule: one thread per pixel, live paths packed between bounces, "
                         "or lanes regenerating paths from a tile queue")
    ap.add_argument("--persistent-blocks", type=int, default=0, help="blocks for the persistent pass (0: fill the GPU)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for {args.gpus}).", file=sys.stderr)
        return 2
    line = bench_raycast(args) if args.workload == RAYCAST else bench_frames(args, args.workload)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
