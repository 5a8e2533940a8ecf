"""The FP64 denoiser the precision variants are measured against
(oracle_denoise64), pinned on the CPU: a uniform image comes back
unchanged, sky pixels average only with sky, and on a noisy image it agrees
with the FP32 restatement of the same passes to FP32 rounding."""
import numpy as np

from oracle.pyoracle import Oracle
from raymond_amd.scenefile import set_params

W, H = 40, 24


def params(total):
    p = np.zeros(208, np.uint8)
    return set_params(p, width=W, height=H, total_samples=total, samples_per_frame=1, denoise=1)


def field(seed, total, sky=False):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:H, 0:W]
    smooth = 0.5 + 0.4 * np.sin(xx / 7.0)[..., None] * np.cos(yy / 5.0)[..., None] * np.array([1.0, 0.8, 0.6])
    albedo = np.clip(0.3 + 0.5 * rng.random((H, W, 3)), 0.0, 1.0)
    light = smooth * (1.0 + 0.3 * rng.standard_normal((H, W, 3)))
    accum = np.zeros((H, W, 4), np.float32)
    accum[..., :3] = light * albedo * total
    accum[..., 3] = (light * albedo).mean(-1) ** 2 * total * 1.1
    galb = np.zeros((H, W, 4), np.float32)
    galb[..., :3] = albedo * total
    n = rng.standard_normal((H, W, 3)) * 0.1 + np.array([0.0, 1.0, 0.0])
    gnorm = np.zeros((H, W, 4), np.float32)
    gnorm[..., :3] = n * total
    gnorm[..., 3] = (3.0 + xx * 0.05) * total
    if sky:
        gnorm[: H // 3] = 0.0
    return accum, galb, gnorm


def test_uniform_image_is_unchanged():
    total = 8
    accum = np.zeros((H, W, 4), np.float32)
    accum[..., :3] = np.array([0.5, 0.25, 0.125]) * total
    accum[..., 3] = 0.1 * total
    galb = np.zeros((H, W, 4), np.float32)
    galb[..., :3] = np.array([0.8, 0.5, 0.25]) * total
    gnorm = np.zeros((H, W, 4), np.float32)
    gnorm[..., 1] = total
    gnorm[..., 3] = 2.0 * total
    out = Oracle.denoise64(params(total), accum, galb, gnorm)
    np.testing.assert_allclose(out[..., :3], accum[..., :3] / total, rtol=1e-12)


def test_sky_pixels_keep_to_the_sky():
    """Rows with no guide (distance 0) are sky: weights 1 among themselves, 0 to surfaces."""
    total = 6
    accum, galb, gnorm = field(3, total, sky=True)
    accum[: H // 3, :, :3] = 0.7 * total      # a flat sky
    galb[: H // 3, :, :3] = total
    out = Oracle.denoise64(params(total), accum, galb, gnorm)
    np.testing.assert_allclose(out[: H // 3 - 2, :, :3], np.float64(accum[0, 0, 0]) / total, rtol=1e-12)


def test_agrees_with_the_fp32_restatement():
    for total, sky in ((2, False), (16, True)):
        accum, galb, gnorm = field(7 + total, total, sky)
        p = params(total)
        from oracle.pyoracle import load, _p
        lib = load()
        a = np.zeros_like(accum)
        b = np.zeros_like(accum)
        lib.oracle_prepare(_p(p), _p(accum), _p(galb), _p(gnorm), _p(a))
        src, dst = a, b
        for step in (1, 2, 4, 8, 16):
            lib.oracle_atrous(_p(p), step, _p(gnorm), _p(src), _p(dst))
            src, dst = dst, src
        out32 = np.zeros_like(accum)
        lib.oracle_compose(_p(p), _p(galb), _p(src), _p(out32))
        out64 = Oracle.denoise64(p, accum, galb, gnorm)
        rel = np.abs(out32[..., :3] - out64[..., :3]) / np.maximum(np.abs(out64[..., :3]), 1e-3)
        assert rel.max() < 2e-4, rel.max()
        assert np.median(rel) < 2e-6
