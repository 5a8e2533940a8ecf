"""Converged accuracy of the plain path on the 200-sphere scene's camera at
160 × 90: 64 frames of 1024 samples each (sample indices 0 .. 65535), the
accumulation read after every frame, so each 1024-sample block mean is a
float64 difference of two reads. The test image is block 0; the reference
is the mean of blocks 1 .. 63, which share no sample index with it.

A pixel is an outlier when its block-0 luminance lies more than 3σ from the
reference, σ² being the spread of the 63 reference block means (which
includes what Owen-scrambled Sobol points gain over independent samples)
plus the reference's own error. How many outliers to expect is measured,
not assumed: each reference block in turn is tested the same way against
the other 62, and their mean outlier rate is the null rate p₀. Under a
normal model p₀ would be P(|t₆₂| > 3) ≈ 0.39%; pixels whose block means
have heavy tails (rare bright paths: caustics, the sun in glossy and glass
reflections) lie outside 3σ more often, and the test prints where the
outliers are. Block 0 must be an ordinary block: its outlier rate at most
p₀ + 4 binomial standard deviations over the image's pixels. Also: relMSE
(tools/converge.mjs:51-61) of block 0 at 1024 spp against the 63-block
reference at most 1e-3, and the sampler no worse than independent sampling
(block spread against the Monte Carlo prediction σ²_sample / 1024 from the
accumulation's Σ luminance², median). With the sun on and the firefly clamp
off (as the bench renders), relMSE is about 1e-2 at 1024 spp and the same
between two 32k-spp halves: a few pixels hold rare samples of the sun
(radiance ≈ 5·10⁴) seen in glossy reflections and through glass, each worth
a large share of its pixel's mean, so the image's relMSE is one draw of a
heavy-tailed sum and does not fall visibly with samples. For both renders
the test checks that the bulk of the image converges at the Monte Carlo
rate — the median per-pixel relMSE term falls at least 8-fold from 1024 spp
(block 0 against the reference) to 32k spp (the two halves against each
other), where independent samples give 16·64/63 — that the worst 1% of
pixels carry most of the half-against-half relMSE when the sun is on, and
that the other 99% meet 1e-3.

Two renders: path tracing alone with the sun off (the sun's caustics are
then rare, huge path-traced samples), and the frame as the bench renders
it, with the sun and light tracing on. There the photon radius is held at
its first frame's value for every block, so that all 64 blocks are
identically distributed (the reference shrinks it frame by frame, which
makes block 0 the most blurred); the gather's blur is then the same bias in
every block and in the reference.
"""
import math

import numpy as np
import pytest

from helpers import rel_mse
from raymond_amd import capi
from raymond_amd.scenefile import get_param, set_params

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]

OPTS = {"scene": {"width": 160, "height": 90}, "width": 160, "height": 90}
BLOCK = 1024
BLOCKS = 64
LUM = np.array([0.2126, 0.7152, 0.0722])


def render_blocks(ctx, p, sun, caustics):
    """(BLOCKS, H, W, 4) per-block means; block b holds sample indices b·BLOCK .. (b+1)·BLOCK - 1."""
    scene = ctx.upload_scene(p)
    r0 = get_param(p.params, "photon_radius")
    try:
        ctx.upload_environment(p)
        ctx.resize(p.width, p.height)
        prev = np.zeros((p.height, p.width, 4), np.float64)
        out = []
        for b in range(BLOCKS):
            off = b * BLOCK
            params = set_params(p.params.copy(), caustics=int(caustics), denoise=0, firefly_clamp=0.0,
                                sun_on=int(sun), samples_per_frame=BLOCK, sample_offset=off, total_samples=off + BLOCK,
                                photon_radius=r0, photon_cell=2 * r0)
            ctx.frame(scene, params, p.emitters if caustics else None)
            cur = ctx.read(capi.BUFFER_ACCUM).astype(np.float64)
            out.append((cur - prev) / BLOCK)
            prev = cur
        return np.stack(out)
    finally:
        scene.release()


def outliers(lum, test, fp32):
    """Pixels where block `test` lies beyond 3σ of the other blocks' mean."""
    ref_blocks = np.delete(lum, test, axis=0)
    ref = ref_blocks.mean(0)
    spread = ref_blocks.var(0, ddof=1)
    sigma = np.sqrt(spread * (1.0 + 1.0 / ref_blocks.shape[0]))
    return np.abs(lum[test] - ref) > 3 * sigma + fp32


@pytest.mark.parametrize("sun,caustics", [(False, False), (True, True)], ids=["paths-only", "as-benched"])
def test_converged_means_within_three_sigma(ctx, pack, sun, caustics):
    p = pack("spheres", **OPTS)
    assert p.meta["causticsActive"]
    blocks = render_blocks(ctx, p, sun, caustics)
    lum = blocks[..., :3] @ LUM                      # block-mean luminance
    # FP32 rounding of the running total: one add per frame, at most half an
    # ulp of the final total each, so at most that much in a block mean.
    peak = blocks[..., :3].max(axis=(0, 3))
    fp32 = np.spacing((peak * BLOCK * BLOCKS).astype(np.float32)).astype(np.float64) / BLOCK
    # Block 0 against blocks 1..63; then the null rate from blocks 1..63,
    # each against the other 62 (block 0 left out of those).
    outside0 = outliers(lum, 0, fp32)
    frac = outside0.mean()
    rest = lum[1:]
    null = np.mean([outliers(rest, b, fp32).mean() for b in range(rest.shape[0])])
    n = outside0.size
    bound = null + 4 * math.sqrt(null * (1 - null) / n)

    # Where the outliers are: the pixels whose block means are most heavy-tailed.
    z = (rest - rest.mean(0)) / np.maximum(rest.std(0, ddof=1), 1e-30)
    kurt = (z ** 4).mean(0) - 3.0
    heavy = kurt > np.quantile(kurt, 0.9)
    share_heavy = outside0[heavy].sum() / max(outside0.sum(), 1)

    def terms(a, r):
        """relMSE's per-pixel terms (tools/converge.mjs:51-61): their mean is rel_mse(a, r)."""
        a, r = a[..., :3], r[..., :3]
        return ((a - r) ** 2 / (r * r + 1e-2)).mean(-1).ravel()

    ref_rgb = blocks[1:].mean(0)
    err = rel_mse(blocks[0].astype(np.float32), ref_rgb.astype(np.float32))
    t_block = terms(blocks[0], ref_rgb)
    half = BLOCKS // 2
    t_half = terms(blocks[:half].mean(0), blocks[half:].mean(0))
    err_half = t_half.mean()
    rate = np.median(t_block) / max(np.median(t_half), 1e-300)
    worst = np.sort(t_half)[::-1]
    k = max(1, t_half.size // 100)
    share_worst = worst[:k].sum() / max(t_half.sum(), 1e-300)
    err_rest = worst[k:].mean()
    ref = rest.mean(0)
    spread = rest.var(0, ddof=1)
    second = blocks[1:, ..., 3].mean(0)              # mean luminance² per sample
    mc = np.maximum(second - ref * ref, 0.0) / BLOCK
    varied = mc > 1e-6 * np.maximum(ref * ref, 1e-6)
    ratio = spread[varied] / mc[varied]
    print(f"\nsun {sun}, light tracing {caustics}: block 0 outside 3 sigma {frac:.4%} (null rate {null:.4%}, bound "
          f"{bound:.4%}; t62 would give 0.39%); {share_heavy:.0%} of block 0's outliers lie in the tenth of pixels "
          f"with the heaviest-tailed block means; relMSE at 1024 spp vs the 63-block reference {err:.3e}, at 32k spp "
          f"vs the other 32k {err_half:.3e} ({share_worst:.0%} of it in the worst 1% of pixels; the rest "
          f"{err_rest:.3e}); median relMSE term 1024 spp / 32k spp {rate:.1f}; block "
          f"spread / MC prediction: median {np.median(ratio):.3f}, p99 {np.percentile(ratio, 99):.3f}")
    assert frac <= bound, (frac, null, bound)
    assert rate >= 8.0, rate
    assert err_rest <= 1e-3, (err_half, share_worst, err_rest)
    if sun:
        assert share_worst > 0.5, (err_half, share_worst)
    else:
        assert err <= 1e-3, err
    assert np.median(ratio) <= 1.0, np.median(ratio)
