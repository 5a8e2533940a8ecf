"""More of the reference's browser tests, restated on the HIP path with their
own thresholds (tests/browser/app.test.mjs):

* shrinking the image and growing it back keeps every pixel in its place
  (:158-174; here bit for bit against the oracle at every size);
* dispersion fringes what is seen through glass, plain glass does not
  (:280-295);
* a patterned lamp sampled directly agrees with the same pattern found by
  bouncing, quarter by quarter of the floor, within 4% (:482-499);
* photon mapping's caustics seen through a glass pane agree with path
  tracing's, within 4% over the floor and 15% in the caustic (:540-569);
* the denoiser: 1 spp × 0.6 and 4 spp × 0.7 of the noisy error (:590-606),
  the checkerboard's edges kept (:608-623), and at 512 spp the denoised
  image within 5% of the brightness of the average, per pixel (:625-634).
"""
import numpy as np
import pytest

from helpers import gpu_render, oracle_render
from test_gpu_closed_form import centre_hits, interior

pytestmark = pytest.mark.gpu


def fixture(pack, name, size, arg=None, bounces=8):
    scene = {"name": name, "size": size, "bounces": bounces}
    if arg is not None:
        scene["arg"] = arg
    return pack("fixture", scene=scene, maxBounces=bounces)


def mean_image(ctx, p, spp, caustics=0, per_frame=64):
    """The average of spp samples: spp / per_frame frames of per_frame samples."""
    per = min(spp, per_frame)
    params = p.with_params(caustics=caustics, denoise=0, firefly_clamp=0.0, samples_per_frame=per)
    return gpu_render(ctx, p, spp // per, params)[0] / spp


def test_resize_keeps_pixels(ctx, pack):
    """One context renders the depth view at 56², 24², 40² and 56² again:
    the per-pixel buffers are kept when the image shrinks and indexed by the
    new width, and every size equals the oracle's depth bit for bit."""
    from raymond_amd import capi
    scenes = {s: pack("mixed", width=s, height=s) for s in (56, 24, 40)}
    scene = ctx.upload_scene(scenes[56])
    try:
        for size in (56, 24, 40, 56):
            p = scenes[size]
            params = p.with_params(caustics=0, denoise=0, view=3)
            ctx.resize(size, size)
            ctx.frame(scene, params)
            gpu = ctx.read(capi.BUFFER_ACCUM)
            cpu = oracle_render(p, 1, params)[0]
            assert (cpu[..., 0] > 0).mean() > 0.2, size   # about a quarter of the view hits an object
            np.testing.assert_array_equal(gpu[..., 0].view(np.uint32), cpu[..., 0].view(np.uint32),
                                          err_msg=f"depth at {size}x{size}")
    finally:
        scene.release()


def colourfulness(img):
    """How far from grey the most colourful pixel is (app.test.mjs:283-287)."""
    rgb = img[..., :3].reshape(-1, 3).astype(np.float64)
    total = rgb.sum(1)
    spread = (rgb.max(1) - rgb.min(1)) / np.where(total > 0.05, total, 1.0)
    return float(np.where(total > 0.05, spread, 0.0).max())


def test_dispersion_fringes(ctx, pack):
    plain = colourfulness(mean_image(ctx, fixture(pack, "stripThroughGlass", 48, "0"), 64))
    split = colourfulness(mean_image(ctx, fixture(pack, "stripThroughGlass", 48, "15"), 64))
    print(f"\ncolourfulness: plain glass {plain:.2e}, splitting glass {split:.3f}")
    assert plain < 1e-3
    assert split > 0.3


def test_patterned_lamp_patch_by_patch(ctx, pack):
    direct = fixture(pack, "patternedLamp", 32, "light", bounces=1)
    bounced = fixture(pack, "patternedLamp", 32, "glow", bounces=1)
    a = mean_image(ctx, direct, 256)
    b = mean_image(ctx, bounced, 1024)
    owner_a, _, _ = centre_hits(direct)
    owner_b, _, _ = centre_hits(bounced)
    floor = interior(owner_a, 0) & (owner_b == 0)
    size = 32
    ys, xs = np.mgrid[0:size, 0:size]
    for qx, qy in ((0, 0), (1, 0), (0, 1), (1, 1)):
        patch = floor & (xs // (size // 2) == qx) & (ys // (size // 2) == qy)
        assert patch.sum() > 20, (qx, qy)
        for c in range(3):
            ra, rb = a[..., c][patch].mean(), b[..., c][patch].mean()
            assert abs(ra / rb - 1.0) < 0.04, f"quarter {(qx, qy)}, channel {c}: direct {ra}, bounced {rb}"


def test_photon_mapping_through_glass(ctx, pack):
    """The floor behind a glass pane, clear of the ball: photon mapping (512
    frames of 1 spp with light tracing, the gather radius shrinking per
    frame) against path tracing alone at 4096 spp."""
    p = fixture(pack, "causticThroughGlass", 32, bounces=10)
    params = p.with_params(caustics=1, denoise=0, firefly_clamp=0.0, samples_per_frame=1)
    mapped = gpu_render(ctx, p, 512, params)[0] / 512
    alone = mean_image(ctx, p, 4096)
    owner, _, rays = centre_hits(p)
    o = rays[:, 0:3].astype(np.float64)
    d = rays[:, 4:7].astype(np.float64)
    oc = o - np.array([0.0, 0.9, 0.0])
    b = (oc * d).sum(1)
    c = (oc * oc).sum(1) - 0.55 * 0.55
    misses_ball = (b * b - (d * d).sum(1) * c < 0).reshape(owner.shape)
    down = (d[:, 1] < 0).reshape(owner.shape)
    pane = interior(owner, 3) & down & misses_ball
    assert pane.sum() > 300, pane.sum()
    bright_m = mapped[..., :3].mean(-1)[pane]
    bright_a = alone[..., :3].mean(-1)[pane]
    assert abs(bright_m.mean() / bright_a.mean() - 1) < 0.04, (bright_m.mean(), bright_a.mean())
    top = np.argsort(-bright_a)[: int(np.ceil(pane.sum() / 4))]
    assert abs(bright_m[top].mean() / bright_a[top].mean() - 1) < 0.15, (bright_m[top].mean(), bright_a[top].mean())


ROOM = ("fixture", {"scene": {"name": "room", "size": 96, "bounces": 4}, "maxBounces": 4})


def rmse(a, b):
    return float(np.sqrt(np.mean((a[..., :3].astype(np.float64) - b[..., :3]) ** 2)))


@pytest.fixture(scope="module")
def room_reference(ctx, pack):
    p = pack(*ROOM[:1], **ROOM[1])
    return p, mean_image(ctx, p, 1024)


@pytest.mark.parametrize("spp,factor", [(1, 0.6), (4, 0.7)])
def test_denoising_brings_few_samples_closer(ctx, room_reference, spp, factor):
    p, ref = room_reference
    params = p.with_params(caustics=0, denoise=1, firefly_clamp=0.0)
    acc, _, _, den = gpu_render(ctx, p, spp, params, denoised=True)
    before, after = rmse(acc / spp, ref), rmse(den, ref)
    print(f"\n{spp} spp: error {before:.4f} noisy, {after:.4f} denoised")
    assert np.isfinite(den).all()
    assert after < before * factor


def test_denoising_keeps_checkerboard_edges(ctx, room_reference):
    """On the checkered floor, blurring the squares together would leave the
    denoised image further from the converged one than the noise was."""
    p, ref = room_reference
    params = p.with_params(caustics=0, denoise=1, firefly_clamp=0.0)
    acc, _, _, den = gpu_render(ctx, p, 4, params, denoised=True)
    owner, _, _ = centre_hits(p)
    floor = interior(owner, 0)
    assert floor.sum() > 1000

    def error(img):
        return float(np.sqrt(np.mean(((img[..., :3].astype(np.float64) - ref[..., :3]) ** 2)[floor])))
    print(f"\nfloor error {error(acc / 4):.4f} noisy, {error(den):.4f} denoised")
    assert error(den) < error(acc / 4)


def test_denoised_becomes_the_average(ctx, pack):
    """At 512 spp the denoised image differs from the plain average by less
    than 5% of its root-mean-square brightness, pixel by pixel (a blurred
    image with the right mean fails this)."""
    p = pack("fixture", scene={"name": "room", "size": 32, "bounces": 4}, maxBounces=4)
    params = p.with_params(caustics=0, denoise=1, firefly_clamp=0.0)
    acc, _, _, den = gpu_render(ctx, p, 512, params, denoised=True)
    avg = acc / 512
    brightness = float(np.sqrt(np.mean(avg[..., :3].astype(np.float64) ** 2)))
    diff = rmse(avg, den)
    print(f"\n512 spp: difference {diff:.4f} against brightness {brightness:.4f}")
    assert diff < 0.05 * brightness
