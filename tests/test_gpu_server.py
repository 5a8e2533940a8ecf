"""The job server (rm_server_*): jobs rendered while others run give the same
image as the same frames rendered alone on a context of their own.

Without light tracing a frame's arithmetic does not depend on scheduling, so
the images agree bit for bit. With light tracing, photons land in the hash
grid in atomic order, so gathered sums agree to float rounding; those jobs
are compared as the light-tracing tests compare against the oracle."""
import time

import numpy as np
import pytest

from helpers import agreement, gpu_render
from raymond_amd import capi

pytestmark = pytest.mark.gpu

JOBS = [
    ("cornell", {"width": 64, "height": 64}, dict(caustics=0, denoise=1, samples_per_frame=2), 6),
    ("spheres", {"scene": {"width": 96, "height": 54}, "width": 96, "height": 54}, dict(caustics=0, denoise=1), 5),
    ("triangles", {"scene": {"grid": 32, "extras": 3, "envWidth": 64, "envHeight": 32, "width": 80, "height": 48},
                   "width": 80, "height": 48}, dict(caustics=0, denoise=1), 4),
    ("fixture", {"scene": {"name": "causticLamp", "size": 48}}, dict(caustics=1, denoise=0), 3),
]


def alone(ctx, p, params, frames):
    out = gpu_render(ctx, p, frames, params, denoised=True)
    return out[3] if params.view(np.uint32)[46] else out[0]


def test_jobs_match_their_frames_rendered_alone(ctx, pack):
    packs = [pack(c, **o) for c, o, _, _ in JOBS]
    params = [p.with_params(firefly_clamp=0.0, **f) for p, (_, _, f, _) in zip(packs, JOBS)]
    frames = [n for *_, n in JOBS]
    expected = [alone(ctx, p, q, n) for p, q, n in zip(packs, params, frames)]
    with capi.Server(slots=3, memory_budget=8 << 30) as server:
        scenes = [server.context.upload_scene(p) for p in packs]
        ids = []
        for _ in range(2):  # every job twice, all in flight together
            for p, q, n, s in zip(packs, params, frames, scenes):
                spp = int(q.view(np.uint32)[43])
                caustics = bool(q.view(np.uint32)[47])
                ids.append(server.submit(s, q, n * spp, p.emitters if caustics else None, packed=p))
        results = [server.wait(i) for i in ids]
    for k, (image, stats) in enumerate(results):
        j = k % len(JOBS)
        name = JOBS[j][0]
        spp = int(params[j].view(np.uint32)[43])
        assert stats["frames"] == frames[j] and stats["samples"] == frames[j] * spp, (name, stats)
        assert 0.0 <= stats["queued_ms"] <= stats["first_frame_ms"] <= stats["done_ms"], (name, stats)
        if params[j].view(np.uint32)[47]:
            frac = agreement(image, expected[j], rel=1e-3, abs_=1e-4)
            assert frac >= 0.999, (name, frac)
        else:
            assert np.array_equal(image.view(np.uint32), expected[j].view(np.uint32)), name


def test_released_buffers_are_reused_safely(ctx, pack):
    """A context's buffers go back to the device's block cache when it closes
    or grows them, and the next context takes them, holding another job's
    values: each job, rendered on a fresh context right after another
    context rendered the previous job and closed, gives the image it gives
    alone."""
    packs = [pack(c, **o) for c, o, _, _ in JOBS]
    params = [p.with_params(firefly_clamp=0.0, **f) for p, (_, _, f, _) in zip(packs, JOBS)]
    frames = [n for *_, n in JOBS]
    expected = [alone(ctx, p, q, n) for p, q, n in zip(packs, params, frames)]
    for j in range(len(JOBS)):
        k = (j + 1) % len(JOBS)
        with capi.Context(0) as before:
            alone(before, packs[k], params[k], frames[k])
        with capi.Context(0) as after:
            image = alone(after, packs[j], params[j], frames[j])
        if params[j].view(np.uint32)[47]:
            assert agreement(image, expected[j], rel=1e-3, abs_=1e-4) >= 0.999, JOBS[j][0]
        else:
            assert np.array_equal(image.view(np.uint32), expected[j].view(np.uint32)), JOBS[j][0]


def test_budget_serialises_jobs(pack):
    """A budget smaller than any job admits one at a time, and they still finish."""
    p = pack("spheres", scene={"width": 64, "height": 36}, width=64, height=36)
    q = p.with_params(caustics=0, denoise=1, firefly_clamp=0.0)
    with capi.Server(slots=4, memory_budget=1) as server:
        scene = server.context.upload_scene(p)
        ids, submitted = [], []
        for _ in range(4):
            t0 = time.perf_counter()
            ids.append(server.submit(scene, q, 3))
            submitted.append((t0, time.perf_counter()))
        results = [server.wait(i) for i in ids]
    first = results[0][0]
    for image, stats in results:
        assert stats["samples"] == 3
        assert np.array_equal(image.view(np.uint32), first.view(np.uint32))
    # Admitted one after another: job k enters its slot only after job k - 1
    # is done. Stats are times since each job's own submit, so job k's
    # admission, measured from job k - 1's submit, is its queued time plus
    # the gap between the two submits, which lies within the host-side
    # bracket of the two calls.
    for k in range(1, len(results)):
        gap_ms = (submitted[k][1] - submitted[k - 1][0]) * 1e3
        admitted_ms = results[k][1]["queued_ms"] + gap_ms
        assert admitted_ms >= results[k - 1][1]["done_ms"], (k, admitted_ms, results[k - 1][1])


def test_precision_options_reach_the_jobs(ctx, pack):
    """rm_server_set_option's precision flags render each job as a context
    with the same flags renders it alone, bit for bit; jobs submitted before
    the call keep the plain path."""
    flags = capi.PRECISION_QUANTIZED_BVH | capi.PRECISION_FP16_GUIDES | capi.PRECISION_COMPENSATED
    chosen = [j for j in JOBS if not j[2].get("caustics")]
    packs = [pack(c, **o) for c, o, _, _ in chosen]
    params = [p.with_params(firefly_clamp=0.0, **f) for p, (_, _, f, _) in zip(packs, chosen)]
    frames = [n for *_, n in chosen]
    plain = [alone(ctx, p, q, n) for p, q, n in zip(packs, params, frames)]
    ctx.set_option(capi.OPTION_PRECISION, flags)
    try:
        varied = [alone(ctx, p, q, n) for p, q, n in zip(packs, params, frames)]
    finally:
        ctx.set_option(capi.OPTION_PRECISION, 0)
    with capi.Server(slots=3, memory_budget=8 << 30) as server:
        scenes = [server.context.upload_scene(p) for p in packs]
        spp = [int(q.view(np.uint32)[43]) for q in params]
        before = [server.submit(s, q, n * k, packed=p) for p, q, n, k, s in zip(packs, params, frames, spp, scenes)]
        server.set_option(capi.OPTION_PRECISION, flags)
        after = [server.submit(s, q, n * k, packed=p) for p, q, n, k, s in zip(packs, params, frames, spp, scenes)]
        with pytest.raises(capi.RaymondError):
            server.set_option(capi.OPTION_PRECISION, 64)
        got_before = [server.wait(i)[0] for i in before]
        got_after = [server.wait(i)[0] for i in after]
    for name, a, b, c, d in zip([j[0] for j in chosen], got_before, plain, got_after, varied):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), name
        assert np.array_equal(c.view(np.uint32), d.view(np.uint32)), name
        assert not np.array_equal(c, a), name


def test_autotuned_jobs_match_their_frames_rendered_alone(ctx, pack):
    """With RM_OPTION_AUTOTUNE the server times the schedules on a job's scene
    first; every schedule gives the same sums, so images stay bit-identical."""
    chosen = [j for j in JOBS if not j[2].get("caustics")]
    packs = [pack(c, **o) for c, o, _, _ in chosen]
    params = [p.with_params(firefly_clamp=0.0, **f) for p, (_, _, f, _) in zip(packs, chosen)]
    frames = [n for *_, n in chosen]
    expected = [alone(ctx, p, q, n) for p, q, n in zip(packs, params, frames)]
    with capi.Server(slots=2, memory_budget=8 << 30) as server:
        server.set_option(capi.OPTION_AUTOTUNE, 1)
        scenes = [server.context.upload_scene(p) for p in packs]
        spp = [int(q.view(np.uint32)[43]) for q in params]
        ids = [server.submit(s, q, n * k, packed=p) for p, q, n, k, s in zip(packs, params, frames, spp, scenes)]
        ids += [server.submit(s, q, n * k, packed=p) for p, q, n, k, s in zip(packs, params, frames, spp, scenes)]
        got = [server.wait(i)[0] for i in ids]
    for k, image in enumerate(got):
        j = k % len(chosen)
        assert np.array_equal(image.view(np.uint32), expected[j].view(np.uint32)), chosen[j][0]


def test_tune_picks_a_schedule_and_keeps_the_image(ctx, pack):
    p = pack("spheres", scene={"width": 96, "height": 54}, width=96, height=54)
    params = p.with_params(caustics=0, denoise=0, firefly_clamp=0.0)
    plain = alone(ctx, p, params, 3)
    scene = ctx.upload_scene(p)
    try:
        ctx.upload_environment(p)
        ctx.resize(p.width, p.height)
        result = ctx.tune(scene, params, frames=2)
    finally:
        scene.release()
    names = [(c[0], c[2], c[3], c[4], c[5]) for c in result["candidates"]]
    assert names == [("split", 4, 8, "squares", 2), ("split", 5, 8, "squares", 2), ("split", 5, 8, "squares", 1),
                     ("per-pixel", 4, 8, "squares", 0), ("per-pixel", 5, 8, "squares", 0),
                     ("per-pixel", 4, 16, "squares", 0), ("per-pixel", 4, 8, "rows", 0),
                     ("compact", 4, 8, "squares", 0), ("sorted", 4, 8, "squares", 0),
                     ("persistent", 4, 8, "squares", 0), ("persistent", 4, 8, "squares", 0),
                     ("wavefront", 4, 8, "squares", 0)]
    assert all(ms > 0 for *_, ms in result["candidates"])
    best = min(result["candidates"], key=lambda c: c[6])
    assert result["best"] == best[:6]
    try:
        tuned = alone(ctx, p, params, 3)
    finally:
        ctx.set_option(capi.OPTION_SCHEDULE, capi.SCHEDULE_DEFAULT)
        ctx.set_option(capi.OPTION_PERSISTENT_BLOCKS, 0)
        ctx.set_option(capi.OPTION_WAVES, 0)
        ctx.set_option(capi.OPTION_STACK_DEPTH, 0)
        ctx.set_option(capi.OPTION_BLOCK_SHAPE, 0)
        ctx.set_option(capi.OPTION_SPLIT_DEPTH, 0)
    assert np.array_equal(tuned.view(np.uint32), plain.view(np.uint32))


def test_tune_tries_only_the_masked_candidates(ctx, pack):
    """RM_OPTION_TUNE_CANDIDATES: the candidates whose bits are set are
    timed, the others list 0 ms, and the pick is among the timed ones."""
    p = pack("spheres", scene={"width": 64, "height": 48}, width=64, height=48)
    params = p.with_params(caustics=0, denoise=0, firefly_clamp=0.0)
    with pytest.raises(capi.RaymondError, match="mask"):
        ctx.set_option(capi.OPTION_TUNE_CANDIDATES, 1 << capi.TUNE_CANDIDATES)
    scene = ctx.upload_scene(p)
    mask = 0b10000000101   # split at four waves, split after one bounce, wavefront
    try:
        ctx.upload_environment(p)
        ctx.resize(p.width, p.height)
        ctx.set_option(capi.OPTION_TUNE_CANDIDATES, mask)
        result = ctx.tune(scene, params, frames=1)
    finally:
        scene.release()
        ctx.set_option(capi.OPTION_TUNE_CANDIDATES, 0)
        ctx.set_option(capi.OPTION_SCHEDULE, capi.SCHEDULE_DEFAULT)
        ctx.set_option(capi.OPTION_PERSISTENT_BLOCKS, 0)
        ctx.set_option(capi.OPTION_WAVES, 0)
        ctx.set_option(capi.OPTION_STACK_DEPTH, 0)
        ctx.set_option(capi.OPTION_BLOCK_SHAPE, 0)
        ctx.set_option(capi.OPTION_SPLIT_DEPTH, 0)
    timed = [k for k, c in enumerate(result["candidates"]) if c[6] > 0]
    assert timed == [k for k in range(capi.TUNE_CANDIDATES) if mask >> k & 1]
    best = min((result["candidates"][k] for k in timed), key=lambda c: c[6])
    assert result["best"] == best[:6]


def test_tune_times_whole_frames_with_light_tracing(ctx, pack):
    """With light tracing and the denoiser on, the tuner renders whole frames
    (the emitters it is given included); without the emitters those frames
    fail."""
    p = pack("fixture", scene={"name": "causticLamp", "size": 48})
    params = p.with_params(caustics=1, denoise=1, firefly_clamp=0.0)
    scene = ctx.upload_scene(p)
    try:
        ctx.upload_environment(p)
        ctx.resize(p.width, p.height)
        ctx.set_option(capi.OPTION_WAVES, 5)
        with pytest.raises(capi.RaymondError):
            ctx.tune(scene, params, frames=1)
        result = ctx.tune(scene, params, frames=1, emitters=p.emitters)
        assert len(result["candidates"]) == capi.TUNE_CANDIDATES
        assert all(ms > 0 for *_, ms in result["candidates"])
        assert result["best"] == min(result["candidates"], key=lambda c: c[6])[:6]
        ctx.frame(scene, params, p.emitters)
        assert np.isfinite(ctx.read(capi.BUFFER_ACCUM)).all()
    finally:
        scene.release()
        ctx.set_option(capi.OPTION_SCHEDULE, capi.SCHEDULE_DEFAULT)
        ctx.set_option(capi.OPTION_PERSISTENT_BLOCKS, 0)
        ctx.set_option(capi.OPTION_WAVES, 0)
        ctx.set_option(capi.OPTION_STACK_DEPTH, 0)
        ctx.set_option(capi.OPTION_BLOCK_SHAPE, 0)
        ctx.set_option(capi.OPTION_SPLIT_DEPTH, 0)


def test_waves_option_is_checked(ctx):
    for bad in (-1, 1, 3, 6):
        with pytest.raises(capi.RaymondError):
            ctx.set_option(capi.OPTION_WAVES, bad)
    for good in (0, 4, 5, 0):
        ctx.set_option(capi.OPTION_WAVES, good)
