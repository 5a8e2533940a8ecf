"""The à-trous denoiser (src/shaders.mjs:1916-2098) against the oracle on the
GPU's own accumulation and guides: prepare's spatial variance estimate (fewer
than 4 samples) and its temporal one, five edge-avoiding passes, compose.
The kernel evaluates each tap's weight with squarings and one exp2 where the
reference calls pow and exp twice, so results agree to a relative 1e-4, the
tolerance stated here. The reference's own denoiser tests (error reduction
at 1 and 4 spp, the checkerboard's edges, convergence to the average pixel
by pixel) are restated in test_gpu_restated.py."""
import numpy as np
import pytest

from helpers import agreement, gpu_render
from oracle.pyoracle import Oracle
from raymond_amd.scenefile import get_param

pytestmark = pytest.mark.gpu

ROOM = ("fixture", {"scene": {"name": "room", "size": 96}})


@pytest.mark.parametrize("frames", [1, 3, 8], ids=["spatial1", "spatial3", "temporal8"])
# 101×67: no side a multiple of the passes' 16-pixel tiles or of their steps,
# so every pass has partial tiles and phases of unequal size; 17×5 and 3×2:
# narrower than the wide steps, so whole phases and workgroups hold no pixel.
@pytest.mark.parametrize("config,opts", [ROOM, ("spheres", {"scene": {"width": 128, "height": 72}, "width": 128,
                                                               "height": 72}),
                                         ("spheres", {"scene": {"width": 101, "height": 67}, "width": 101,
                                                      "height": 67}),
                                         ("spheres", {"scene": {"width": 17, "height": 5}, "width": 17, "height": 5}),
                                         ("spheres", {"scene": {"width": 3, "height": 2}, "width": 3, "height": 2})],
                         ids=["room", "spheres", "spheres101x67", "spheres17x5", "spheres3x2"])
def test_denoiser_matches_oracle(ctx, pack, config, opts, frames):
    from helpers import frame_params
    p = pack(config, **opts)
    params = p.with_params(caustics=0, denoise=1, firefly_clamp=0.0)
    acc, ga, gn, den = gpu_render(ctx, p, frames, params, denoised=True)
    last = frame_params(p, frames - 1, params)
    assert get_param(last, "total_samples") == frames
    passes = Oracle(p).denoise(last, acc, ga, gn)
    ref = passes[-1]
    frac = agreement(den, ref, rel=1e-4, abs_=1e-6)
    assert frac > 0.995, f"{frac:.4f} of pixels agree"
    assert np.isfinite(den).all()
