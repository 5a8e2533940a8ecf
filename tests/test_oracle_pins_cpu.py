"""Closed forms that pin the oracle's denoiser, GGX and display arithmetic
(test infrastructure checks of oracle/oracle.cpp; the GPU is held to the
oracle elsewhere).

Denoiser (src/shaders.mjs:1916-2098), with guides chosen so that each
weight factor can be isolated:
* one à-trous pass's impulse response is the outer product of the B3
  spline (1/16, 1/4, 3/8, 1/4, 1/16) dilated by the step, and the variance
  of an impulse propagates as Σw²v/(Σw)²;
* across a crease at angle θ a neighbour counts cos⁶⁴θ, and across a depth
  step |Δz| it counts exp(−|Δz|/(0.02·z·step + 1e-6)), which gives a
  closed-form output beside the edge;
* across a luminance step ΔL a neighbour counts exp(−ΔL/σ) with
  σ = 4√variance + 1e-4;
* prepare's variance from n ≥ 4 samples is the unbiased sample variance of
  the luminance over n, in units of the light (divided by albedo²), and
  from fewer it is the neighbourhood's spread plus 0.01·mean²;
* the FP64 denoiser agrees with the FP32 one on random inputs.

GGX (Walter et al. 2007, with α = max(roughness², 0.002)): D against the
textbook α²/(π cos⁴θ (α² + tan²θ)²), normalised (∫D cosθ dω = 1 by
quadrature); Smith G1 against 2/(1 + √(1 + α² tan²θ)); the sampler's half
vectors against the closed-form CDF tan²θ/(α² + tan²θ); sample_bsdf's mean
weight against the quadrature of eval_bsdf (a wrong pdf or Jacobian fails).

Display: Narkowicz's ACES fit and the sRGB transfer function (IEC
61966-2-1) at known points.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle
from oracle.pyoracle import Oracle
from raymond_amd.scenefile import set_params

K = np.array([0.0625, 0.25, 0.375, 0.25, 0.0625])   # B3 spline weights at offsets -2..2


def params(w, h, n=1):
    return set_params(np.zeros(208, np.uint8), width=w, height=h, total_samples=n, samples_per_frame=1)


def guides(h, w, n=1, normal=(0.0, 0.0, 1.0), depth=1.0):
    g = np.zeros((h, w, 4), np.float32)
    g[..., :3] = np.asarray(normal, np.float32) * n
    g[..., 3] = depth * n
    return g


def atrous(p, step, gn, light):
    out = np.zeros_like(light)
    lib = pyoracle.load()
    lib.oracle_atrous(pyoracle._p(p), step, pyoracle._p(gn), pyoracle._p(light), pyoracle._p(out))
    return out


@pytest.mark.parametrize("step", [1, 2, 4])
def test_atrous_impulse_response_is_the_dilated_b3_spline(step):
    """Uniform guides and a variance so large that σ makes every luminance
    weight 1: the pass is the plain dilated B3 filter, exactly."""
    h = w = 41
    p = params(w, h)
    light = np.zeros((h, w, 4), np.float32)
    light[..., 3] = 1e30
    cy = cx = 20
    light[cy, cx, :3] = 1.0
    out = atrous(p, step, guides(h, w), light)
    want = np.zeros((h, w))
    for i, dy in enumerate(range(-2, 3)):
        for j, dx in enumerate(range(-2, 3)):
            want[cy + dy * step, cx + dx * step] = K[i] * K[j]
    np.testing.assert_array_equal(out[..., 0], want.astype(np.float32))
    np.testing.assert_array_equal(out[..., 1], out[..., 0])
    # Variance: Σw²v / (Σw)² with Σw = 1 and v = 1e30 everywhere.
    np.testing.assert_allclose(out[cy, cx, 3], 1e30 * (K ** 2).sum() ** 2, rtol=1e-6)


def edge_output(c):
    """Beside a vertical edge (the pixel just left of it), with value 0 on
    the left and 1 on the right and every right-hand tap weighted c more:
    taps at dx = +1, +2 hold B3 weight 1/4 + 1/16, the rest 11/16."""
    return 0.3125 * c / (0.6875 + 0.3125 * c)


def two_halves(h, w, left, right):
    a = np.zeros((h, w, 4), np.float32)
    a[:, : w // 2] = left
    a[:, w // 2:] = right
    return a


@pytest.mark.parametrize("theta", [0.05, 0.1, 0.2])
def test_normal_weight_is_cos_to_the_64th(theta):
    h, w = 9, 16
    p = params(w, h)
    gn = two_halves(h, w, [0, 0, 1, 1], [math.sin(theta), 0, math.cos(theta), 1])
    light = two_halves(h, w, [0, 0, 0, 1e30], [1, 1, 1, 1e30])
    out = atrous(p, 1, gn, light)
    c = np.float32(math.cos(theta)) ** 64
    np.testing.assert_allclose(out[4, w // 2 - 1, 0], edge_output(float(c)), rtol=2e-5)


@pytest.mark.parametrize("step", [1, 2])
def test_depth_weight_closed_form(step):
    h, w = 9, 24
    p = params(w, h)
    z1, z2 = 2.0, 2.1
    gn = two_halves(h, w, [0, 0, 1, z1], [0, 0, 1, z2])
    light = two_halves(h, w, [0, 0, 0, 1e30], [1, 1, 1, 1e30])
    out = atrous(p, step, gn, light)
    c = math.exp(-abs(z1 - z2) / (z1 * 0.02 * step + 1e-6))
    # The pixel just left of the edge: its taps at dx = +1, +2 (s and 2s
    # pixels away) lie across it, the rest on its own side, at any step.
    np.testing.assert_allclose(out[4, w // 2 - 1, 0], edge_output(c), rtol=1e-5)


@pytest.mark.parametrize("variance", [0.01, 0.25])
def test_luminance_weight_uses_four_sigma(variance):
    """A grey step of luminance 1 with uniform guides and uniform variance v:
    σ = 4√v + 1e-4 (the 3×3-blurred variance of a constant is itself)."""
    h, w = 9, 16
    p = params(w, h)
    light = two_halves(h, w, [0, 0, 0, variance], [1, 1, 1, variance])
    out = atrous(p, 1, guides(h, w), light)
    # Grey of 1 has luminance (0.2126 + 0.7152) + 0.0722 in FP32.
    lum = float(np.float32(np.float32(0.2126) + np.float32(0.7152)) + np.float32(0.0722))
    sigma = 4.0 * math.sqrt(variance) + 1e-4
    c = math.exp(-lum / sigma)
    np.testing.assert_allclose(out[4, w // 2 - 1, 0], edge_output(c), rtol=1e-5)


def prepare(p, acc, ga, gn):
    out = np.zeros_like(acc)
    lib = pyoracle.load()
    lib.oracle_prepare(pyoracle._p(p), pyoracle._p(acc), pyoracle._p(ga), pyoracle._p(gn), pyoracle._p(out))
    return out


def test_prepare_variance_from_the_samples():
    """Eight grey samples of brightness l on albedo 0.5: the light is the
    mean over the albedo, and the variance the unbiased sample variance of
    the luminance, over n, over the albedo's luminance squared."""
    h, w, n = 4, 4, 8
    samples = np.array([0.5, 1.0, 1.5, 2.0, 0.5, 1.0, 1.5, 2.0])
    p = params(w, h, n)
    acc = np.zeros((h, w, 4), np.float32)
    acc[..., :3] = samples.sum()
    acc[..., 3] = (samples ** 2).sum()
    ga = np.full((h, w, 4), 0.5 * n, np.float32)
    gn = guides(h, w, n)
    out = prepare(p, acc, ga, gn)
    a = 0.5   # luminance of grey 0.5 (the weights sum to 1 within an ulp)
    var = samples.var(ddof=1) / n / (a * a)
    np.testing.assert_allclose(out[..., :3], samples.mean() / 0.5, rtol=1e-5)
    np.testing.assert_allclose(out[..., 3], var, rtol=1e-4)


def test_prepare_variance_from_the_neighbourhood():
    """Three samples (fewer than four): a constant image has no spread, so
    its variance is 0.01·mean²; a pixel with zero albedo guide counts as
    albedo 0.01."""
    h, w, n = 6, 6, 3
    p = params(w, h, n)
    acc = np.zeros((h, w, 4), np.float32)
    acc[..., :3] = 0.75 * n
    ga = np.full((h, w, 4), 0.5 * n, np.float32)
    out = prepare(p, acc, ga, guides(h, w, n))
    np.testing.assert_allclose(out[..., :3], 1.5, rtol=1e-6)
    np.testing.assert_allclose(out[..., 3], 0.01 * 1.5 ** 2, rtol=1e-5)
    ga[2, 3] = 0.0
    out = prepare(p, acc, ga, guides(h, w, n))
    np.testing.assert_allclose(out[2, 3, :3], 0.75 / 0.01, rtol=1e-6)


def test_fp64_denoiser_agrees_with_fp32():
    rng = np.random.default_rng(9)
    h, w, n = 40, 56, 5
    p = params(w, h, n)
    acc = rng.uniform(0, 2, (h, w, 4)).astype(np.float32) * n
    acc[..., 3] = (acc[..., :3].mean(-1) / n) ** 2 * n * 1.3
    ga = rng.uniform(0.1, 0.9, (h, w, 4)).astype(np.float32) * n
    gn = np.zeros((h, w, 4), np.float32)
    nrm = rng.normal(0, 0.2, (h, w, 3)) + [0, 0, 1]
    gn[..., :3] = nrm / np.linalg.norm(nrm, axis=-1, keepdims=True) * n
    gn[..., 3] = rng.uniform(1.0, 1.2, (h, w)) * n

    out32 = Oracle.denoise(type("O", (), {"lib": pyoracle.load()})(), p, acc, ga, gn)[-1]
    out64 = Oracle.denoise64(p, acc, ga, gn)
    np.testing.assert_allclose(out32[..., :3], out64[..., :3], rtol=2e-4, atol=1e-6)


# ---------------------------------------------------------------- GGX

def ggx_textbook_d(cos_t, alpha):
    tan2 = (1 - cos_t ** 2) / cos_t ** 2
    return alpha ** 2 / (np.pi * cos_t ** 4 * (alpha ** 2 + tan2) ** 2)


@pytest.mark.parametrize("roughness", [0.01, 0.2236, 0.5477, 0.8944])
def test_ggx_d_is_the_textbook_distribution(roughness):
    """α = max(r², 0.002): r = 0.01 hits the floor; the others give α of
    0.05, 0.3 and 0.8."""
    alpha = max(roughness ** 2, 0.002)
    for c in np.linspace(0.05, 1.0, 40):
        d = pyoracle.ggx_terms(roughness, c, 1.0, 1.0)[0]
        # FP32 evaluates k = cos²θ(α² − 1) + 1, which cancels to about α²
        # near cosθ = 1: a few ulps of 1 relative to k, squared in D.
        k = c * c * (alpha ** 2 - 1) + 1
        np.testing.assert_allclose(d, ggx_textbook_d(c, alpha), rtol=8 * 2.0 ** -24 / k + 2e-5)


@pytest.mark.parametrize("roughness", [0.2236, 0.5477, 0.8944])
def test_ggx_d_is_normalised(roughness):
    """∫ D(h) cosθ_h dω = 1 over the hemisphere, by quadrature in cosθ."""
    from scipy.integrate import quad
    f = lambda c: float(pyoracle.ggx_terms(roughness, c, 1.0, 1.0)[0]) * c * 2 * np.pi
    total = quad(f, 0.0, 1.0, points=[0.9, 0.99, 0.999], limit=400, epsabs=1e-6)[0]
    assert abs(total - 1.0) < 1e-3, total


@pytest.mark.parametrize("roughness", [0.2236, 0.5477, 0.8944])
def test_smith_g1_is_the_textbook_form(roughness):
    alpha = roughness ** 2
    for c in np.linspace(0.05, 1.0, 30):
        g = pyoracle.ggx_terms(roughness, 1.0, c, c)[1]
        tan2 = (1 - c * c) / (c * c)
        np.testing.assert_allclose(g, 2 / (1 + math.sqrt(1 + alpha ** 2 * tan2)), rtol=2e-5)


@pytest.mark.parametrize("roughness", [0.2236, 0.5477])
def test_ggx_sampler_follows_d_cos(roughness):
    """Half vectors from a 256×256 grid of uniforms: their θ has the CDF
    tan²θ/(α² + tan²θ) (within 0.003, a Kolmogorov distance far below what
    a wrong α or exponent gives) and φ is uniform."""
    alpha = roughness ** 2
    u = (np.arange(256) + 0.5) / 256
    hs = np.array([pyoracle.ggx_sample(roughness, a, b) for a in u[::1] for b in u[::8]])
    np.testing.assert_allclose(np.linalg.norm(hs, axis=1), 1, atol=1e-6)
    cos_t = np.sort(np.clip(hs[:, 2], 0, 1))[::-1]
    tan2 = (1 - cos_t ** 2) / np.maximum(cos_t ** 2, 1e-12)
    cdf = tan2 / (alpha ** 2 + tan2)
    emp = (np.arange(len(cos_t)) + 0.5) / len(cos_t)
    assert np.max(np.abs(cdf - emp)) < 0.003
    phi = np.arctan2(hs[:, 1], hs[:, 0])
    hist = np.histogram(phi, bins=8, range=(-np.pi, np.pi))[0]
    assert hist.min() > 0.9 * hist.mean()


def material(kind, color=(1.0, 1.0, 1.0), roughness=0.3, ior=1.5):
    w = np.zeros(24, np.uint32)
    f = w.view(np.float32)
    f[0:3] = color
    w[3] = kind
    f[7] = roughness
    f[12] = ior
    w[13] = w[16] = w[18] = 0xFFFFFFFF
    return w


def hemisphere(n_t=192, n_p=192):
    t = (np.arange(n_t) + 0.5) / n_t * (np.pi / 2)
    ph = (np.arange(n_p) + 0.5) / n_p * 2 * np.pi
    T, P = np.meshgrid(t, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    dw = (np.sin(T) * (np.pi / 2 / n_t) * (2 * np.pi / n_p)).ravel()
    return d.astype(np.float32), dw


@pytest.mark.parametrize("kind,roughness", [(0, 0.5), (1, 0.5), (1, 0.8), (2, 0.3), (2, 0.7)],
                         ids=["diffuse", "metal0.5", "metal0.8", "plastic0.3", "plastic0.7"])
def test_sample_bsdf_mean_weight_is_the_albedo(kind, roughness):
    """E[f·cos/pdf] over sample_bsdf's draws equals ∫ f·cos dω by quadrature
    of eval_bsdf: the sampler's pdf matches the density it draws from."""
    n = np.array([0, 0, 1.0], np.float32)
    wo = np.array([0.5, 0.0, math.sqrt(0.75)], np.float32)
    m = material(kind, roughness=roughness)
    dirs, dw = hemisphere()
    albedo = sum(pyoracle.eval_bsdf(m, [1, 1, 1], n, wo, dirs[i])[0] * dw[i] for i in range(len(dirs)))
    draws = np.array([pyoracle.sample_bsdf(m, [1, 1, 1], n, wo, seed) for seed in range(40000)])
    weights = np.where(draws[:, 8] > 0, draws[:, 3], 0.0)
    mean, err = weights.mean(), weights.std() / math.sqrt(len(weights))
    assert abs(mean - albedo) < 4 * err + 2e-3, (mean, err, albedo)


# ---------------------------------------------------------------- display

def aces_narkowicz(x):
    """Narkowicz (2015), "ACES Filmic Tone Mapping Curve": the published fit."""
    return np.clip(x * (2.51 * x + 0.03) / (x * (2.43 * x + 0.59) + 0.14), 0.0, 1.0)


def srgb_encode(x):
    """IEC 61966-2-1: linear below 0.0031308, else 1.055·x^(1/2.4) − 0.055."""
    x = np.clip(x, 0.0, 1.0)
    return np.where(x <= 0.0031308, 12.92 * x, 1.055 * x ** (1 / 2.4) - 0.055)


def test_display_is_aces_then_srgb():
    values = np.array([0.0, 0.001, 0.01, 0.05, 0.18, 0.5, 1.0, 2.0, 4.0, 16.0], np.float32)
    acc = np.zeros((1, len(values), 4), np.float32)
    acc[0, :, :3] = values[:, None] * 4
    p = set_params(params(len(values), 1, 4), exposure=1.0, view=0, denoise=0)
    shown = pyoracle.display(p, acc)[0, :, 0].astype(int)
    want = np.round(srgb_encode(aces_narkowicz(values.astype(np.float64))) * 255).astype(int)
    assert np.abs(shown - want).max() <= 1, (shown, want)
    # Known points: middle grey 0.18 → ACES 0.2669 → sRGB 0.5534 → 141;
    # 1.0 → ACES 2.54/3.16 = 0.8038 → sRGB 0.9082 → 232.
    assert int(shown[4]) == 141 and int(shown[6]) == 232, shown
