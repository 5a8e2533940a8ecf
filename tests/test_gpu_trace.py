"""Bit-exact traversal parity: the HIP closest-hit and any-hit kernels against
the oracle's FP32 restatement of trace() (src/shaders.mjs:389-447). Every
traced ray must give the same primitive index and the same t, u and v bits."""
import numpy as np
import pytest

from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


def random_rays(n, seed, spread=15.0, target=5.0):
    rng = np.random.default_rng(seed)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = rng.uniform(-spread, spread, (n, 3))
    aim = rng.uniform(-target, target, (n, 3))
    d = aim - rays[:, 0:3]
    # Some unnormalised directions, some axis-aligned ones (safe() paths).
    d[: n // 8] *= rng.uniform(0.1, 10.0, (n // 8, 1))
    d[n // 8: n // 8 + n // 32, 1:] = 0.0
    rays[:, 4:7] = d
    rays[:, 3] = 1e30
    rays[n // 2: n // 2 + n // 16, 3] = rng.uniform(0.5, 20.0, n // 16)
    return rays


def assert_same_hits(gpu, cpu, what):
    gpu = gpu.reshape(-1, 4)
    cpu = cpu.reshape(-1, 4)
    bad = np.nonzero(np.any(gpu != cpu, axis=1))[0]
    assert bad.size == 0, (
        f"{what}: {bad.size} of {len(cpu)} hit records differ; first {bad[:5].tolist()}: "
        f"gpu {gpu[bad[:3]].tolist()} cpu {cpu[bad[:3]].tolist()}")


@pytest.mark.parametrize("any_hit", [False, True])
def test_cornell_primary_rays(ctx, pack, any_hit):
    p = pack("cornell")
    o = Oracle(p)
    rays = o.camera_rays()
    assert rays.shape == (256 * 256, 8)
    scene = ctx.upload_scene(p)
    try:
        gpu = ctx.raycast(scene, rays, any_hit=any_hit)
    finally:
        scene.release()
    cpu = o.trace(rays, any_hit=any_hit)
    assert (cpu[:, 1] != 0xffffffff).mean() > 0.9
    assert_same_hits(gpu, cpu, "cornell primary rays")


@pytest.mark.parametrize("config,seed", [("mixed", 101), ("spheres", 202), ("cornell", 303)])
def test_random_rays(ctx, pack, config, seed):
    p = pack(config)
    o = Oracle(p)
    rays = random_rays(200_000, seed=seed)
    scene = ctx.upload_scene(p)
    try:
        for any_hit in (False, True):
            gpu = ctx.raycast(scene, rays, any_hit=any_hit)
            cpu = o.trace(rays, any_hit=any_hit)
            assert_same_hits(gpu, cpu, f"{config} random rays any_hit={any_hit}")
    finally:
        scene.release()


def test_traversal_counters_match(ctx, pack):
    p = pack("mixed")
    o = Oracle(p)
    rays = random_rays(50_000, seed=3)
    scene = ctx.upload_scene(p)
    try:
        _, gcnt = ctx.raycast(scene, rays, counters=True)
    finally:
        scene.release()
    _, ccnt = o.trace(rays, counters=True)
    np.testing.assert_array_equal(gcnt, ccnt)
