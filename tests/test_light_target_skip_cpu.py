"""Light tracing counts the aim targets whose cone holds a light path's
direction (src/shaders.mjs's emission loop; light.hip light_paths). The kernel
skips a target, without its divisions, where the test is sure to fail:
s = d·to < 0 with s² > 1e-8 d2, or s >= 0 with s² < 0.999 (d2 - r²) - 1e-4 d2.
Here the kernel's FP32 test and the skip are run in float32 (numpy's division
and square root are correctly rounded, as the kernels are compiled; no FMA),
on directions crowding the cone's edge and the plane through the origin, and
no skipped target may pass the test."""
import numpy as np

F = np.float32


def dot(a, b):
    return (a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]) + a[..., 2] * b[..., 2]


def normalize(a):
    return a / np.sqrt(dot(a, a))[..., None]


def kernel_test(d, to, r):
    """1 - d·(to/|to|) <= sphere_omc(r, |to|²), as the kernel computes it."""
    d2 = dot(to, to)
    x = np.minimum(r * r / d2, F(1))
    omc = x / (F(1) + np.sqrt(F(1) - x))
    return F(1) - dot(d, to / np.sqrt(d2)[..., None]) <= omc


def kernel_skip(d, to, r):
    s, d2 = dot(d, to), dot(to, to)
    return np.where(s < F(0), s * s > F(1e-8) * d2, s * s < F(0.999) * (d2 - r * r) - F(1e-4) * d2)


def test_skipped_targets_never_pass_the_cone_test():
    rng = np.random.default_rng(7)
    n = 600_000
    to = normalize(rng.uniform(-1, 1, (n, 3))) * (10.0 ** rng.uniform(-2, 2, n))[:, None]
    to = to.astype(F)
    d2 = dot(to, to).astype(np.float64)
    # radii from tiny to just inside the distance
    frac = np.where(rng.random(n) < 0.25, 1 - 10.0 ** rng.uniform(-7, -1, n), rng.random(n))
    r = (np.sqrt(d2) * frac).astype(F)
    keep = dot(to, to) > r * r
    cos_a = np.sqrt(np.maximum(0.0, 1 - r.astype(np.float64) ** 2 / d2))
    mode = np.arange(n) % 3
    jitter = (rng.uniform(-1, 1, n)) * 10.0 ** rng.uniform(-7, -1, n)
    theta = np.where(mode == 0, np.arccos(np.minimum(cos_a, 1)) * (1 + jitter),
                     np.where(mode == 1, np.pi / 2 + jitter, np.arccos(rng.uniform(-1, 1, n))))
    phi = rng.uniform(0, 2 * np.pi, n)
    a = to.astype(np.float64) / np.sqrt(d2)[:, None]
    helper = np.where(np.abs(a[:, :1]) < 0.9, [[1.0, 0, 0]], [[0, 1.0, 0]])
    t = np.cross(a, helper)
    t /= np.linalg.norm(t, axis=1)[:, None]
    b = np.cross(a, t)
    d = (a * np.cos(theta)[:, None] + t * (np.sin(theta) * np.cos(phi))[:, None]
         + b * (np.sin(theta) * np.sin(phi))[:, None]).astype(F)
    d = normalize(d)
    d, to, r = d[keep], to[keep], r[keep]
    passes, skipped = kernel_test(d, to, r), kernel_skip(d, to, r)
    assert not np.any(passes & skipped), f"{int(np.sum(passes & skipped))} skipped targets pass the test"
    # the skip does its job: most of these targets are skipped, and the test still passes on many
    assert skipped.mean() > 0.4 and passes.mean() > 0.15
