"""Failure paths on the GPU: arguments refused before anything is queued
leave the context rendering as a fresh one does; samples whose radiance or
guides are not finite are dropped, guides and all, as the reference drops
them (src/shaders.mjs:1875-1883); served jobs can be cancelled or miss a
deadline; running out of device memory is an error, in Python and in JS."""
import os
import subprocess
import time
from pathlib import Path

import numpy as np
import pytest

from helpers import agreement, gpu_render, oracle_render
from raymond_amd import capi

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent

LAMP = ("fixture", {"scene": {"name": "causticLamp", "size": 48}})


def test_bad_emitters_are_refused_and_leave_the_context_intact(ctx, pack):
    """Light tracing needs a target to aim at and quad or sphere sources: an
    emitter table without targets, or naming a triangle, is refused before
    anything is queued, and the next frames render as on a fresh context."""
    p = pack(*LAMP[:1], **LAMP[1])
    params = p.with_params(caustics=1, denoise=1, firefly_clamp=0.0)
    no_targets = p.emitters.copy()
    no_targets.view(np.uint32)[1] = 0
    scene = ctx.upload_scene(p)
    try:
        ctx.upload_environment(p)
        ctx.resize(p.width, p.height)
        with pytest.raises(capi.RaymondError, match="at least one emitter and one target") as e:
            ctx.frame(scene, params, no_targets)
        assert e.value.code == capi.ERR_INVALID
    finally:
        scene.release()
    tri = pack("caustics", scene={"res": 8, "width": 32, "height": 18}, width=32, height=18)
    triangles = np.nonzero(tri.primitives.reshape(-1, 16)[:, 0] == 3)[0]
    as_triangle = tri.emitters.copy()
    as_triangle.view(np.uint32)[4 + 1] = int(triangles[0])   # the first source's primitive
    tparams = tri.with_params(caustics=1, denoise=0, firefly_clamp=0.0)
    scene = ctx.upload_scene(tri)
    try:
        ctx.resize(tri.width, tri.height)
        with pytest.raises(capi.RaymondError, match="neither a quad nor a sphere"):
            ctx.frame(scene, tparams, as_triangle)
    finally:
        scene.release()
    after = gpu_render(ctx, p, 3, params, denoised=True)
    with capi.Context(0) as fresh:
        clean = gpu_render(fresh, p, 3, params, denoised=True)
    for a, b in zip(after, clean):
        assert agreement(a, b, rel=1e-5, abs_=1e-6) == 1.0


def non_finite_sky(params):
    """The parameter block with a sky that is not finite: horizon +Inf,
    zenith NaN, ground -Inf (words 16-18, 20-22, 24-26)."""
    out = params.copy()
    f = out.view(np.float32)
    f[16:19] = np.nan
    f[20:23] = np.inf
    f[24:27] = -np.inf
    return out


def test_non_finite_samples_are_dropped(ctx, pack):
    """The Cornell box is open at +z: paths that leave through the opening
    meet a sky that is not finite. With one sample per pixel every pixel is
    either exactly the clean render's (its path never left) or zero in its
    sums and guides (its sample dropped); the oracle drops the same ones, and
    the denoised image stays finite."""
    p = pack("cornell", width=64, height=64)
    clean_params = p.with_params(caustics=0, denoise=1, firefly_clamp=0.0)
    bad_params = non_finite_sky(clean_params)
    clean = gpu_render(ctx, p, 1, clean_params)
    bad = gpu_render(ctx, p, 1, bad_params, denoised=True)
    for name, b in zip(("accum", "albedo", "normal", "denoised"), bad):
        assert np.isfinite(b).all(), name
    sums = np.concatenate(bad[:3], axis=-1)
    kept = (sums.view(np.uint32) == np.concatenate(clean, axis=-1).view(np.uint32)).all(axis=-1)
    dropped = (sums == 0.0).all(axis=-1)
    assert (kept | dropped).all()
    print(f"\nnon-finite sky: {dropped.mean():.4f} of samples dropped, {kept.mean():.4f} kept")
    assert 0.01 < dropped.mean() < 0.9
    cpu = oracle_render(p, 1, bad_params)
    for name, g, c in zip(("accum", "albedo", "normal"), bad, cpu):
        assert agreement(g, c) == 1.0, name
    # Later frames keep adding only finite samples; the denoiser stays finite.
    more = gpu_render(ctx, p, 4, bad_params, denoised=True)
    assert all(np.isfinite(b).all() for b in more)


def test_cancel_and_deadline(ctx, pack):
    """One slot: a long job is cancelled while it runs, a queued one before
    it starts, a third misses a 1 ms deadline; each wait raises
    RM_ERR_CANCELLED. A job submitted after them renders its image exactly
    as it does alone."""
    p = pack("spheres", scene={"width": 128, "height": 72}, width=128, height=72)
    params = p.with_params(caustics=0, denoise=1, firefly_clamp=0.0)
    expected = gpu_render(ctx, p, 3, params, denoised=True)[3]
    with capi.Server(slots=1, memory_budget=8 << 30) as server:
        scene = server.context.upload_scene(p)
        long_job = server.submit(scene, params, 100_000)
        queued = server.submit(scene, params, 100_000)
        late = server.submit(scene, params, 100_000, deadline_ms=1.0)
        good = server.submit(scene, params, 3)
        server.cancel(queued)
        time.sleep(0.2)
        server.cancel(long_job)
        outcomes = {}
        for name, job in (("running", long_job), ("queued", queued), ("late", late)):
            with pytest.raises(capi.RaymondError) as e:
                server.wait(job)
            assert e.value.code == capi.ERR_CANCELLED, (name, str(e.value))
            outcomes[name] = str(e.value)
        image, stats = server.wait(good)
        with pytest.raises(capi.RaymondError):
            server.cancel(good)   # waited for: no longer known
    assert "cancelled" in outcomes["running"] and "cancelled" in outcomes["queued"]
    assert "deadline" in outcomes["late"]
    assert stats["samples"] == 3
    assert np.array_equal(image.view(np.uint32), expected.view(np.uint32))


def test_out_of_memory_is_an_error(ctx, pack):
    """Two 16383² images hold more than the GPU's 288 GB (per-pixel buffers
    and traversal spill, about 167 GB each): the second context's resize
    fails with RM_ERR_MEMORY. The failure is reported once: frames on the
    failed context, on the first and on any other render as before."""
    p = pack("cornell", width=32, height=32)
    params = p.with_params(caustics=0, denoise=1, firefly_clamp=0.0)
    expected = gpu_render(ctx, p, 2, params, denoised=True)
    with capi.Context(0) as a, capi.Context(0) as b:
        a.resize(16383, 16383)
        with pytest.raises(capi.RaymondError, match="out of memory") as e:
            b.resize(16383, 16383)
        assert e.value.code == capi.ERR_MEMORY, str(e.value)
        for c in (b, a, ctx):
            got = gpu_render(c, p, 2, params, denoised=True)
            for g, x in zip(got, expected):
                assert np.array_equal(g.view(np.uint32), x.view(np.uint32))


def test_errors_reach_javascript():
    """Through the addon and the JS Renderer: running out of device memory,
    and a refused argument, are thrown as Errors carrying the library's
    sentence; a cancelled served job's wait() rejects."""
    res = subprocess.run(["node", str(ROOT / "tests" / "js" / "errors_gpu.mjs")], capture_output=True, text=True,
                         timeout=300, env=dict(os.environ, RAYMOND_EXPECT_GPU="1"))
    assert res.returncode == 0, res.stdout + res.stderr
    print("\n" + res.stdout.strip())
