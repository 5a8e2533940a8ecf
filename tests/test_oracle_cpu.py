"""The oracle pinned on the CPU: the reference's own known-answer raycasts
(tests/unit/scene.test.mjs:143-315), brute force over every primitive, FP32
against FP64, and closed forms and invariants for the RNG, materials,
dispersion, the environment's sampling density and the denoiser."""
import math

import numpy as np
import pytest

from oracle import pyoracle
from oracle.pyoracle import Oracle
from raymond_amd.scenefile import get_param

WHITE = {"type": "diffuse", "color": [0.8, 0.8, 0.8]}
NONE = 0xFFFFFFFF


def ray(o, d, t_max=1e30):
    return np.array([*o, t_max, *d, 0.0], np.float32)


def first_hit(p, o, d, t_max=1e30):
    o32 = Oracle(p)
    h = o32.trace(ray(o, d, t_max)[None])[0]
    h64 = o32.trace64(ray(o, d, t_max)[None])[0]
    t = h[0:1].view(np.float32)[0]
    return (None if h[1] == NONE else float(t)), (None if h64[1] < 0 else float(h64[0])), h


def scene(pack, objects, **kw):
    return pack("custom", scene=dict(objects=objects, **kw))


def test_sphere_near_side_and_inside(pack):
    p = scene(pack, [{"type": "sphere", "center": [0, 0, 0], "radius": 1, "material": WHITE}])
    t, t64, _ = first_hit(p, (0, 0, 5), (0, 0, -1))
    assert t == 4.0 and abs(t64 - 4) < 1e-12
    t, t64, _ = first_hit(p, (0, 0, 0), (1, 0, 0))
    assert t == 1.0 and abs(t64 - 1) < 1e-12
    assert first_hit(p, (0, 2, 5), (0, 0, -1))[0] is None
    assert first_hit(p, (0, 0, 5), (0, 0, -1), 3.9)[0] is None, "beyond the maximum distance"


def test_distance_in_units_of_direction(pack):
    p = scene(pack, [{"type": "sphere", "center": [0, 0, 0], "radius": 1, "material": WHITE}])
    t, t64, _ = first_hit(p, (0, 0, 5), (0, 0, -2))
    assert abs(t - 2) < 1e-6 and abs(t64 - 2) < 1e-12


def test_quad_inside_edges_from_either_side(pack):
    p = scene(pack, [{"type": "quad", "corner": [-1, 0, 1], "u": [2, 0, 0], "v": [0, 0, -2], "material": WHITE}])
    assert abs(first_hit(p, (0.5, 3, 0.5), (0, -1, 0))[0] - 3) < 1e-6
    assert abs(first_hit(p, (0.5, -3, 0.5), (0, 1, 0))[0] - 3) < 1e-6
    assert first_hit(p, (1.5, 3, 0), (0, -1, 0))[0] is None


def test_turned_stretched_box(pack):
    # translate(0,1,0) · rotateY(π/2) · scale(2,1,1), row-major.
    c, s = math.cos(math.pi / 2), math.sin(math.pi / 2)
    m = np.array([[c, 0, s, 0], [0, 1, 0, 1], [-s, 0, c, 0], [0, 0, 0, 1]]) @ np.diag([2, 1, 1, 1])
    p = scene(pack, [{"type": "box", "transform": m.ravel().tolist(), "material": WHITE}])
    for o, d, t in [((0, 1, 10), (0, 0, -1), 8), ((10, 1, 0), (-1, 0, 0), 9), ((0.5, 10, 1.5), (0, -1, 0), 8),
                    ((0, 1, 0), (0, 1, 0), 1)]:
        got, got64, _ = first_hit(p, o, d)
        assert abs(got - t) < 1e-5 and abs(got64 - t) < 1e-5, (o, got, got64)


def test_stick_between_points(pack):
    a, b = np.array([1, 2, 3.0]), np.array([-2, 4, 0.0])
    along = (b - a) / np.linalg.norm(b - a)
    p = scene(pack, [{"type": "stick", "a": a.tolist(), "b": b.tolist(), "thickness": 0.2, "material": WHITE}])
    got, got64, _ = first_hit(p, tuple(a - along), tuple(along))
    assert abs(got - 1.0) < 1e-5 and abs(got64 - 1.0) < 1e-5
    q = scene(pack, [{"type": "stick", "a": [0, 0, 0], "b": [0, 3, 0], "thickness": 0.1, "material": WHITE}])
    assert abs(first_hit(q, (0, 10, 0), (0, -1, 0))[0] - 7) < 1e-5


def random_rays(n, seed):
    rng = np.random.default_rng(seed)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = rng.uniform(-15, 15, (n, 3))
    r[:, 4:7] = rng.uniform(-0.5, 0.5, (n, 3))
    r[:, 3] = 1e30
    return r


def test_nearest_of_many_matches_brute_force(pack):
    """The BVH's closest hit equals the minimum over every primitive's own
    test (tests/unit/scene.test.mjs:233-276), on 20,000 random rays."""
    p = pack("mixed")
    o = Oracle(p)
    rays = random_rays(20000, 11)
    hits = o.trace(rays)
    brute = o.brute_force(rays)
    assert np.array_equal(hits[:, 1] == NONE, brute[:, 1] == NONE)
    m = hits[:, 1] != NONE
    assert np.array_equal(hits[m, 0], brute[m, 0]), "closest t differs from exhaustive search"
    # Equal t from two primitives (a shared edge) may name either; otherwise the same one.
    same = hits[m, 1] == brute[m, 1]
    assert same.mean() > 0.999
    checked = int(m.sum())
    assert checked > 50


def test_fp32_agrees_with_fp64(pack):
    p = pack("mixed")
    o = Oracle(p)
    rays = random_rays(100000, 5)
    h32, h64 = o.trace(rays), o.trace64(rays)
    prim32 = np.where(h32[:, 1] == NONE, -1, h32[:, 1].astype(np.int64))
    agree = (prim32 == h64[:, 1].astype(np.int64)).mean()
    assert agree > 0.999
    hit = (prim32 >= 0) & (prim32 == h64[:, 1])
    rel = np.abs(h32[hit, 0].view(np.float32) - h64[hit, 0]) / np.abs(h64[hit, 0])
    assert np.median(rel) < 1e-6 and np.quantile(rel, 0.999) < 1e-3


def test_any_hit_finds_a_hit_when_closest_does(pack):
    p = pack("spheres", scene={"width": 64, "height": 36}, width=64, height=36)
    o = Oracle(p)
    rays = random_rays(20000, 2)
    closest, anyh = o.trace(rays), o.trace(rays, any_hit=True)
    assert np.array_equal(closest[:, 1] == NONE, anyh[:, 1] == NONE)
    # Any hit's t is never nearer than the closest.
    m = closest[:, 1] != NONE
    assert (anyh[m, 0].view(np.float32) >= closest[m, 0].view(np.float32)).all()


def test_empty_scene_hits_nothing(pack):
    p = scene(pack, [])
    assert p.primitive_count == 0
    assert first_hit(p, (0, 0, 5), (0, 0, -1))[0] is None


# ---------------------------------------------------------------- RNG

def test_sobol2_is_linear_and_a_net():
    """(reverseBits(i), sobol2(i)) for i < 2^m is a (0, m, 2)-net: every
    elementary box of area 2^-m holds exactly one point."""
    for a, b in [(3, 5), (1234, 4321), (0xFFFF, 1)]:
        assert pyoracle.sobol2(a ^ b) == pyoracle.sobol2(a) ^ pyoracle.sobol2(b)
    for m in range(1, 11):
        pts = [(pyoracle.reverse_bits(i), pyoracle.sobol2(i)) for i in range(1 << m)]
        for k in range(m + 1):
            cells = {(x >> (32 - k) if k else 0, y >> (32 - (m - k)) if m - k else 0) for x, y in pts}
            assert len(cells) == 1 << m, (m, k)


def test_sobol2_superset_transform():
    """The kernels compute sobol2 as a superset XOR transform over bit
    positions followed by a bit reversal (the generator matrix is Pascal's
    triangle mod 2); it equals the oracle's bit walk."""
    rng = np.random.default_rng(5)
    xs = np.concatenate([rng.integers(0, 1 << 32, 1 << 14, dtype=np.uint64),
                         np.array([0, 1, 2, 3, 0x80000000, 0xFFFFFFFF, 0x55555555, 0xAAAAAAAA], np.uint64)])
    w = xs.copy()
    for shift, mask in [(1, 0x55555555), (2, 0x33333333), (4, 0x0F0F0F0F), (8, 0x00FF00FF), (16, 0x0000FFFF)]:
        w ^= (w >> np.uint64(shift)) & np.uint64(mask)
    for x, v in zip(xs.tolist(), w.tolist()):
        assert pyoracle.sobol2(x) == int(f"{v:032b}"[::-1], 2), hex(x)


def test_owen_scrambled_pairs_keep_the_net():
    """A pixel's first 256 rand2() pairs in one dimension stay a (0, 8, 2)-net
    after Owen scrambling (src/shaders.mjs:161-212)."""
    for pixel, dim in [(0, 0), (12345, 1 << 16), (777, (3 << 16) + 2)]:
        pts = np.array([pyoracle.qmc_draws(pixel, s, dim, 1, True)[0] for s in range(256)])
        xi = np.floor(pts * (1 << 24)).astype(np.int64)
        for k in range(9):
            cells = {(x >> (24 - k), y >> (24 - (8 - k))) for x, y in xi}
            assert len(cells) == 256, (pixel, dim, k)


def test_dimensions_are_independent():
    a = np.array([pyoracle.qmc_draws(p, 7, 1 << 16, 1, False)[0] for p in range(4096)])
    b = np.array([pyoracle.qmc_draws(p, 7, 2 << 16, 1, False)[0] for p in range(4096)])
    assert abs(np.corrcoef(a, b)[0, 1]) < 0.05
    assert abs(a.mean() - 0.5) < 0.02 and a.max() < 1.0 and a.min() >= 0.0


def test_pcg_is_uniform_and_avalanches():
    vals = np.array([pyoracle.pcg(i) for i in range(200000)], np.uint64)
    counts = np.bincount((vals >> 24).astype(np.int64), minlength=256)
    expected = len(vals) / 256
    chi2 = ((counts - expected) ** 2 / expected).sum()
    assert chi2 < 350, chi2  # 255 degrees of freedom
    flips = [bin(pyoracle.pcg(i) ^ pyoracle.pcg(i ^ (1 << b))).count("1") for i in range(0, 4000, 7) for b in range(32)]
    assert abs(np.mean(flips) - 16) < 0.5


# ---------------------------------------------------------------- materials

def material(kind, color=(0.8, 0.8, 0.8), roughness=0.3, ior=1.5):
    w = np.zeros(24, np.uint32)
    f = w.view(np.float32)
    f[0:3] = color
    w[3] = kind
    f[7] = roughness
    f[12] = ior
    w[13] = w[16] = w[18] = NONE
    return w


def hemisphere(n_t=128, n_p=256):
    t = (np.arange(n_t) + 0.5) / n_t * (np.pi / 2)
    ph = (np.arange(n_p) + 0.5) / n_p * 2 * np.pi
    T, P = np.meshgrid(t, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    dw = (np.sin(T) * (np.pi / 2 / n_t) * (2 * np.pi / n_p)).ravel()
    return d.astype(np.float32), dw


def test_diffuse_is_lambert():
    n = np.array([0, 0, 1.0], np.float32)
    wo = np.array([0.3, 0.2, 0.9], np.float32) / np.linalg.norm([0.3, 0.2, 0.9])
    wi = np.array([-0.5, 0.1, 0.7], np.float32) / np.linalg.norm([-0.5, 0.1, 0.7])
    e = pyoracle.eval_bsdf(material(0), [0.8, 0.5, 0.2], n, wo, wi)
    c = wi[2]
    np.testing.assert_allclose(e[:3], np.array([0.8, 0.5, 0.2]) * c / np.pi, rtol=1e-6)
    np.testing.assert_allclose(e[3], c / np.pi, rtol=1e-6)


@pytest.mark.parametrize("kind,rough", [(1, 0.3), (1, 0.6), (2, 0.2), (2, 0.5)])
def test_glossy_pdf_and_energy(kind, rough):
    """The sampling pdf integrates to at most 1 over the hemisphere (GGX
    reflections below the horizon are lost), and a white surface reflects no
    more than it receives."""
    n = np.array([0, 0, 1.0], np.float32)
    wo = np.array([0.4, 0.0, 0.916515], np.float32)
    dirs, dw = hemisphere()
    pdf = np.zeros(len(dirs))
    f = np.zeros(len(dirs))
    m = material(kind, (1, 1, 1), rough)
    for i in range(0, len(dirs)):
        e = pyoracle.eval_bsdf(m, [1, 1, 1], n, wo, dirs[i])
        pdf[i], f[i] = e[3], e[0]
    total = (pdf * dw).sum()
    assert 0.8 < total <= 1.01, total
    assert (f * dw).sum() <= 1.01


def test_glossy_reciprocity():
    n = np.array([0, 0, 1.0], np.float32)
    a = np.array([0.3, 0.4, 0.866], np.float32)
    b = np.array([-0.5, 0.2, 0.842615], np.float32)
    for kind in (1, 2):
        m = material(kind, (0.9, 0.6, 0.3), 0.4)
        fab = pyoracle.eval_bsdf(m, [0.9, 0.6, 0.3], n, a, b)
        fba = pyoracle.eval_bsdf(m, [0.9, 0.6, 0.3], n, b, a)
        if kind == 1:  # metal: f·cos/cos symmetric
            np.testing.assert_allclose(fab[:3] / b[2], fba[:3] / a[2], rtol=1e-4)


def test_fresnel_closed_forms():
    for eta in (1.33, 1.5, 2.4):
        assert abs(pyoracle.fresnel_dielectric(1.0, eta) - ((eta - 1) / (eta + 1)) ** 2) < 1e-6
    assert pyoracle.fresnel_dielectric(0.1, 1 / 1.5) == 1.0, "total internal reflection"
    assert pyoracle.fresnel_dielectric(1e-4, 1.5) > 0.99, "grazing"


def test_wavelength_colours_average_white():
    """Wyman-Sloan-Shirley CIE fit, clipped and scaled so the average over 380-720 nm is white."""
    ls = np.linspace(380, 720, 3401)
    rgb = np.array([pyoracle.wavelength_rgb(l) for l in ls])
    np.testing.assert_allclose(rgb.mean(0), [1, 1, 1], atol=0.02)
    assert (rgb >= 0).all()


def test_environment_pdf_integrates_to_one(pack):
    """tests/unit/environment.test.mjs:87-114, by quadrature over the sphere."""
    p = pack("fixture", scene={"name": "environmentFloor", "size": 8})
    o = Oracle(p)
    import ctypes as C
    n_t, n_p = 128, 256
    total = 0.0
    for i in range(n_t):
        th = (i + 0.5) / n_t * np.pi
        for j in range(n_p):
            ph = (j + 0.5) / n_p * 2 * np.pi
            d = np.array([np.sin(th) * np.cos(ph), np.cos(th), np.sin(th) * np.sin(ph)], np.float32)
            total += o.lib.oracle_env_pdf(o.ref, p.params.ctypes.data_as(C.c_void_p),
                                          d.ctypes.data_as(C.c_void_p)) * np.sin(th)
    total *= (np.pi / n_t) * (2 * np.pi / n_p)
    assert abs(total - 1.0) < 0.02, total


def test_denoiser_keeps_a_constant_image(pack):
    """With one colour, one albedo and one normal everywhere, every weight is
    normalised away: prepare, five à-trous passes and compose return the input."""
    p = pack("fixture", scene={"name": "room", "size": 24})
    h, w = p.height, p.width
    n = 5
    params = p.with_params(total_samples=n)
    acc = np.zeros((h, w, 4), np.float32)
    acc[..., :3] = np.array([0.3, 0.6, 0.9]) * n
    acc[..., 3] = (0.2126 * 0.3 + 0.7152 * 0.6 + 0.0722 * 0.9) ** 2 * n
    ga = np.zeros_like(acc)
    ga[..., :3] = np.array([0.5, 0.5, 0.5]) * n
    gn = np.zeros_like(acc)
    gn[..., 1] = n
    gn[..., 3] = 2.0 * n
    out = Oracle(p).denoise(params, acc, ga, gn)[-1]
    np.testing.assert_allclose(out[..., :3], np.broadcast_to([0.3, 0.6, 0.9], (h, w, 3)), rtol=1e-5)


def test_camera_rays_follow_the_basis(pack):
    """Pixel-centre rays: unit length, the centre pixel's along forward, corners at ±tan(fov/2)."""
    p = pack("cornell", width=64, height=64)
    rays = Oracle(p).camera_rays().reshape(64, 64, 8)
    d = rays[..., 4:7]
    np.testing.assert_allclose(np.linalg.norm(d, axis=-1), 1, atol=1e-6)
    assert np.allclose(rays[..., 0:3], [0, 1, 3.85])
    t = get_param(p.params, "tan_half_fov")
    assert abs(t - math.tan(37 * math.pi / 360)) < 1e-6
    mid = (d[31, 31] + d[32, 32]) / 2
    assert abs(mid[2] + 1) < 1e-3
    x = d[0, 0, 0] / -d[0, 0, 2]
    assert abs(x - (-(1 - 1 / 64) * t)) < 1e-5


def test_display_closed_forms():
    """DISPLAY_SHADER (src/shaders.mjs:2100-2152) restated: ACES is 0 at 0 and
    saturates, sRGB meets its linear segment at 0.0031308 and maps 1 to 1,
    normals map [-1, 1] to [0, 1], depth halves at the focus distance, the
    accumulation is averaged first, and nothing rendered shows black."""
    from raymond_amd import scenefile
    from raymond_amd.scenefile import get_param, set_params
    p = scenefile.pack("cornell", {"width": 8, "height": 1})
    acc = np.zeros((1, 8, 4), np.float32)
    values = [0.0, 0.0031308, 1.0, 1e6]
    for i, v in enumerate(values):
        acc[0, i, :3] = v * 4  # four samples
    view = lambda k, **kw: pyoracle.display(set_params(p.params.copy(), total_samples=4, denoise=0, view=k, **kw), acc)
    albedo = view(1)[0]
    assert list(albedo[0, :3]) == [0, 0, 0]
    # Both sRGB branches give 0.04045 at the threshold: round(0.04045 * 255) = 10.
    assert list(albedo[1, :3]) == [10, 10, 10]
    assert list(albedo[2, :3]) == [255, 255, 255]
    colour = view(0, exposure=1.0)[0]
    assert list(colour[0, :3]) == [0, 0, 0] and list(colour[3, :3]) == [255, 255, 255]
    ramp = np.zeros((1, 8, 4), np.float32)
    ramp[0, :, :3] = np.linspace(0.0, 4.0, 8)[:, None] * 4
    shades = pyoracle.display(set_params(p.params.copy(), total_samples=4, denoise=0, view=0, exposure=1.0), ramp)
    assert np.all(np.diff(shades[0, :, 0].astype(int)) >= 0), "ACES then sRGB is monotone"
    normals = np.zeros((1, 8, 4), np.float32)
    normals[0, 0, :3] = np.array([-1.0, 0.0, 1.0]) * 4
    assert list(pyoracle.display(set_params(p.params.copy(), total_samples=4, view=2), normals)[0, 0, :3]) == [0, 128, 255]
    focus = float(get_param(p.params, "focus_distance"))
    depth = np.zeros((1, 8, 4), np.float32)
    depth[0, 0, 0] = focus * 4
    assert list(pyoracle.display(set_params(p.params.copy(), total_samples=4, view=3), depth)[0, :2, 0]) == [128, 0]
    assert np.all(pyoracle.display(set_params(p.params.copy(), total_samples=0), acc)[0] == [0, 0, 0, 255])
