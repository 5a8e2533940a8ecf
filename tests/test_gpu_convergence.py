"""Converged accuracy of the plain path on the 200-sphere scene's camera at
160 × 90, light tracing and the sun off (the 0.5° sun's radiance is about
5·10⁴; its caustics through the glass balls are rare, huge samples that path
tracing alone cannot converge at 1024 spp; with light tracing on, as the
bench renders, those paths come from splats instead).

64 frames of 1024 samples each (sample indices 0 .. 65535) are rendered and
the accumulation is read after every frame, so each 1024-sample block mean
is a float64 difference of two reads: a one-frame register sum per pixel,
then a single FP32 add into the running total. The test image is block 0;
the reference is the mean of blocks 1 .. 63, which share no sample index
with it.

A pixel passes when its block-0 luminance lies within 3σ of the reference,
where σ² is the spread of the 63 reference block means (which includes what
Owen-scrambled Sobol points gain over independent samples) plus the
reference's own error. Under a normal model 0.27% of pixels fall outside 3σ
by chance; the bound asserted is 1%. Two more checks: the image's relMSE
(tools/converge.mjs:51-61) is at most 1e-3, and the sampler is no worse than
independent sampling: the block means' spread is at most the Monte Carlo
prediction σ²_sample / 1024 (from the accumulation's Σ luminance²) for the
median pixel, and not far above it for the 99th percentile.
"""
import numpy as np
import pytest

from helpers import rel_mse
from raymond_amd import capi
from raymond_amd.scenefile import set_params

pytestmark = pytest.mark.gpu

OPTS = {"scene": {"width": 160, "height": 90}, "width": 160, "height": 90}
BLOCK = 1024
BLOCKS = 64
LUM = np.array([0.2126, 0.7152, 0.0722])


def render_blocks(ctx, p):
    """(BLOCKS, H, W, 4) per-block means; block b holds sample indices b·BLOCK .. (b+1)·BLOCK - 1."""
    scene = ctx.upload_scene(p)
    try:
        ctx.upload_environment(p)
        ctx.resize(p.width, p.height)
        prev = np.zeros((p.height, p.width, 4), np.float64)
        out = []
        for b in range(BLOCKS):
            off = b * BLOCK
            ctx.frame(scene, set_params(p.params.copy(), caustics=0, denoise=0, firefly_clamp=0.0, sun_on=0,
                                        samples_per_frame=BLOCK, sample_offset=off, total_samples=off + BLOCK))
            cur = ctx.read(capi.BUFFER_ACCUM).astype(np.float64)
            out.append((cur - prev) / BLOCK)
            prev = cur
        return np.stack(out)
    finally:
        scene.release()


def test_converged_means_within_three_sigma(ctx, pack):
    p = pack("spheres", **OPTS)
    blocks = render_blocks(ctx, p)
    lum = blocks[..., :3] @ LUM                      # block-mean luminance
    test, ref_blocks = lum[0], lum[1:]
    ref = ref_blocks.mean(0)
    n_ref = ref_blocks.shape[0]
    spread = ref_blocks.var(0, ddof=1)               # variance of one block's mean
    sigma = np.sqrt(spread * (1.0 + 1.0 / n_ref))    # of (test - ref)
    # FP32 rounding of the running total: one add per frame, at most half an
    # ulp of the final total each, so at most that much in a block mean.
    peak = blocks[..., :3].max(axis=(0, 3))
    fp32 = np.spacing((peak * BLOCK * BLOCKS).astype(np.float32)).astype(np.float64) / BLOCK
    outside = np.abs(test - ref) > 3 * sigma + fp32
    frac = outside.mean()

    ref_rgb = blocks[1:].mean(0)
    err = rel_mse(blocks[0].astype(np.float32), ref_rgb.astype(np.float32))

    second = blocks[1:, ..., 3].mean(0)              # mean luminance² per sample
    mc = np.maximum(second - ref * ref, 0.0) / BLOCK
    varied = mc > 1e-6 * np.maximum(ref * ref, 1e-6)
    ratio = spread[varied] / mc[varied]
    print(f"outside 3 sigma: {frac:.4%}; relMSE at 1024 spp vs 64k-spp reference: {err:.3e}; "
          f"block spread / MC prediction: median {np.median(ratio):.3f}, p99 {np.percentile(ratio, 99):.3f}")
    assert frac < 0.01, frac
    assert err <= 1e-3, err
    assert np.median(ratio) <= 1.0, np.median(ratio)
    assert np.percentile(ratio, 99) <= 2.0, np.percentile(ratio, 99)
