"""The display pass (rm_display) against the oracle's DISPLAY_SHADER
(src/shaders.mjs:2100-2152) on the same accumulation, in every view: RGBA8
equal up to one step where powf's last ulp tips a rounding."""
import numpy as np
import pytest

from helpers import frame_params
from oracle import pyoracle
from raymond_amd import capi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("view", [0, 1, 2, 3], ids=["color", "albedo", "normal", "depth"])
def test_display_matches_oracle(ctx, pack, view):
    p = pack("cornell", width=64, height=48)
    params = p.with_params(caustics=0, denoise=1 if view == 0 else 0, firefly_clamp=0.0, view=view, exposure=1.5)
    scene = ctx.upload_scene(p)
    try:
        ctx.upload_environment(p)
        ctx.resize(p.width, p.height)
        frames = 3
        for k in range(frames):
            ctx.frame(scene, frame_params(p, k, params))
        shown = ctx.display()
        accum = ctx.read(capi.BUFFER_ACCUM)
        denoised = ctx.read(capi.BUFFER_DENOISED) if view == 0 else None
    finally:
        scene.release()
    expect = pyoracle.display(frame_params(p, frames - 1, params), accum, denoised)
    diff = np.abs(shown.astype(int) - expect.astype(int))
    assert diff.max() <= 1, diff.max()
    assert (diff == 0).mean() > 0.99
    assert (shown[..., 3] == 255).all()
    assert shown[..., :3].max() > 0, "the image isn't black"
