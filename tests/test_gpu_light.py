"""Light tracing, splats and photon mapping against the oracle.

A frame with caustics on runs the trace pass (which leaves caustic paths to
light tracing), light_trace (photons through mirrors and glass, splatted to
the camera in 1/65536 fixed point and deposited in the hash grid),
resolve_splats and photon_gather (src/shaders.mjs:1425-1831). Photon paths use
the same PCG streams on both sides; splats are integer sums, so they agree
exactly wherever the paths do. Photons land in the grid in a different order
on the GPU (one counter atomic per wave), so gathered sums agree to float
rounding rather than bit for bit."""
import numpy as np
import pytest

from helpers import agreement, gpu_render, oracle_render

pytestmark = pytest.mark.gpu

CAUSTIC_SCENES = [
    ("fixture", {"scene": {"name": "causticLamp", "size": 64}}),
    ("fixture", {"scene": {"name": "causticSun", "size": 64}}),
    ("fixture", {"scene": {"name": "causticMirror", "size": 64}}),
    ("fixture", {"scene": {"name": "causticThroughGlass", "size": 64}}),
    ("cornell", {"width": 96, "height": 96}),
    ("caustics", {"scene": {"res": 24}, "width": 96, "height": 54}),
]
IDS = ["lamp", "sun", "mirror", "through_glass", "cornell", "dispersive_water"]


@pytest.mark.parametrize("config,opts", CAUSTIC_SCENES, ids=IDS)
def test_caustic_frame_matches_oracle(ctx, pack, config, opts):
    p = pack(config, **opts)
    assert p.meta["causticsActive"], "the scene should turn light tracing on"
    params = p.with_params(caustics=1, denoise=0, firefly_clamp=0.0)
    g = gpu_render(ctx, p, 2, params)[0]
    c = oracle_render(p, 2, params)[0]
    assert np.isfinite(g).all()
    frac = agreement(g, c, rel=1e-3, abs_=1e-4)
    print(f"\n{config} {opts}: {frac:.5f} of pixels agree")
    assert frac >= 0.999, f"only {frac:.4f} of pixels agree"
    total_g, total_c = g[..., :3].sum(), c[..., :3].sum()
    assert abs(total_g - total_c) <= 0.01 * abs(total_c) + 1e-3, (total_g, total_c)


def test_light_tracing_adds_light(ctx, pack):
    """With caustics on, the caustic under the glass ball comes from splats and gathers, not from nothing."""
    p = pack(*CAUSTIC_SCENES[0][:1], **CAUSTIC_SCENES[0][1])
    on = gpu_render(ctx, p, 4, p.with_params(caustics=1, denoise=0, firefly_clamp=0.0))[0]
    off_params = p.with_params(caustics=1, denoise=0, firefly_clamp=0.0)
    # The trace pass alone with caustics set leaves the caustic paths out.
    only_trace = oracle_render(p, 4, off_params, light=False)[0]
    assert on[..., :3].sum() > only_trace[..., :3].sum() * 1.01


@pytest.mark.parametrize("config,opts", [CAUSTIC_SCENES[0], CAUSTIC_SCENES[5]], ids=["lamp", "dispersive_water"])
def test_denoised_frames_resolve_the_same_sums(ctx, pack, config, opts):
    """From the fourth sample on, a denoised frame resolves its splats and
    gathered light inside the denoiser's first pass (k_finish) rather than in
    k_resolve: the sums equal those of frames without the denoiser, to the
    rounding of the gathered sums (photon order), and the splats alone bit
    for bit (light tracing's PCG streams are the same)."""
    p = pack(config, **opts)
    on = gpu_render(ctx, p, 6, p.with_params(caustics=1, denoise=1, firefly_clamp=0.0))[0]
    off = gpu_render(ctx, p, 6, p.with_params(caustics=1, denoise=0, firefly_clamp=0.0))[0]
    assert np.isfinite(on).all()
    np.testing.assert_allclose(on, off, rtol=1e-5, atol=1e-6)


def test_switching_scenes_on_one_context(pack):
    """Light-traced frames of one scene, then of another with other emitters
    and a larger image, on one context: the second scene's sums equal those it
    gets on a fresh context, to the rounding of the gathered sums. (The
    emitters are uploaded again when they change, and growing the image
    clears both splat buffers.)"""
    from raymond_amd import capi
    a = pack(*CAUSTIC_SCENES[1][:1], **CAUSTIC_SCENES[1][1])
    b = pack("caustics", scene={"res": 24, "width": 160, "height": 90}, width=160, height=90)
    pa = a.with_params(caustics=1, denoise=1, firefly_clamp=0.0)
    pb = b.with_params(caustics=1, denoise=1, firefly_clamp=0.0)
    with capi.Context(0) as one:
        gpu_render(one, a, 5, pa)
        switched = gpu_render(one, b, 5, pb)[0]
    with capi.Context(0) as fresh:
        alone = gpu_render(fresh, b, 5, pb)[0]
    assert np.isfinite(switched).all()
    np.testing.assert_allclose(switched, alone, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("paths", [3 * 64 * 64, (1 << 20) + 12345], ids=["3x-pixels", "past-the-grid"])
def test_light_path_budget_matches_oracle(ctx, pack, paths):
    """RM_OPTION_LIGHT_PATHS: that many light paths a frame (and photon
    slots), normalised by their number, against the oracle with the same
    budget. 2^20 + 12345 paths pass the kernel's grid of 2^20 threads, so
    threads take several photon ids each."""
    from raymond_amd import capi
    p = pack(*CAUSTIC_SCENES[0][:1], **CAUSTIC_SCENES[0][1])
    params = p.with_params(caustics=1, denoise=0, firefly_clamp=0.0)
    ctx.set_option(capi.OPTION_LIGHT_PATHS, paths)
    try:
        g = gpu_render(ctx, p, 2, params)[0]
    finally:
        ctx.set_option(capi.OPTION_LIGHT_PATHS, 0)
    c = oracle_render(p, 2, params, paths=paths)[0]
    plain = gpu_render(ctx, p, 2, params)[0]
    frac = agreement(g, c, rel=1e-3, abs_=1e-4)
    print(f"\n{paths} light paths a frame: {frac:.5f} of pixels agree")
    assert frac >= 0.999
    assert abs(g[..., :3].sum() - c[..., :3].sum()) <= 1e-3 * abs(c[..., :3].sum())
    assert not np.array_equal(g, plain), "a budget past the reference's count traces more photons"


def test_light_path_budget_converges_like_path_tracing(ctx, pack):
    """64 times the reference's light paths a frame over 128 frames, normalised by their
    number: the floor still converges to what path tracing alone finds
    (whole floor within 4%, brightest decile within 8%, the reference's
    thresholds)."""
    from raymond_amd import capi
    from test_gpu_closed_form import average, centre_hits, fixture, interior
    p = fixture(pack, "causticLamp", 48)
    ctx.set_option(capi.OPTION_LIGHT_PATHS, 64 * 48 * 48)
    try:
        on = gpu_render(ctx, p, 128, p.with_params(caustics=1, denoise=0, firefly_clamp=0.0))[0] / 128
    finally:
        ctx.set_option(capi.OPTION_LIGHT_PATHS, 0)
    off = average(ctx, p, 16384, caustics=0)
    owner, _, _ = centre_hits(p)
    mask = interior(owner, 0)
    a, b = on[mask][:, :3].mean(1), off[mask][:, :3].mean(1)
    assert abs(a.mean() / b.mean() - 1) < 0.04, (a.mean(), b.mean())
    top = b >= np.quantile(b, 0.9)
    assert abs(a[top].mean() / b[top].mean() - 1) < 0.08, (a[top].mean(), b[top].mean())
