"""The path-tracing kernel against the oracle under identical RNG streams:
one sample per pixel, so each pixel's accumulation is one radiance() value.
Paths only part where a transcendental (cos, exp, pow, ...) differs by an ulp
between the GPU's and the CPU's maths library and flips a branch, so almost
every pixel must agree to float rounding."""
import numpy as np
import pytest

from oracle.pyoracle import Oracle
from raymond_amd import capi

pytestmark = pytest.mark.gpu


def render_first_frame(ctx, p, params):
    scene = ctx.upload_scene(p)
    try:
        ctx.resize(p.width, p.height)
        ctx.frame(scene, params)
        return ctx.read(capi.BUFFER_ACCUM), ctx.read(capi.BUFFER_GUIDE_ALBEDO), ctx.read(capi.BUFFER_GUIDE_NORMAL)
    finally:
        scene.release()


def oracle_first_frame(p, params):
    o = Oracle(p)
    shape = (p.height, p.width, 4)
    acc, ga, gn = (np.zeros(shape, np.float32) for _ in range(3))
    o.main(params, acc, ga, gn)
    return acc, ga, gn


def agreement(a, b, rel=1e-4, abs_=1e-5):
    close = np.abs(a - b) <= abs_ + rel * np.abs(b)
    return close.all(axis=-1).mean()


@pytest.mark.parametrize("config,opts", [
    ("spheres", {"scene": {"width": 160, "height": 90}, "width": 160, "height": 90}),
    ("cornell", {"width": 96, "height": 96}),
])
def test_first_frame_per_sample(ctx, pack, config, opts):
    p = pack(config, **opts)
    params = p.with_params(caustics=0, denoise=0, firefly_clamp=0.0)
    gpu = render_first_frame(ctx, p, params)
    cpu = oracle_first_frame(p, params)
    for name, g, c in zip(("accum", "albedo", "normal"), gpu, cpu):
        assert np.isfinite(g).all(), name
        frac = agreement(g, c)
        assert frac > 0.97, f"{config} {name}: only {frac:.4f} of pixels agree"
