"""The HIP traversal against the committed golden hit records
(tests/golden/make_golden.py), through the C ABI: bit-exact for closest and
any-hit, with the same traversal counters."""
import numpy as np
import pytest

from helpers import load_golden

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["cornell_c1.npz", "spheres_c2.npz"])
def test_raycast_matches_golden(ctx, name):
    p, rays, closest, any_hit, counters = load_golden(name)
    scene = ctx.upload_scene(p)
    try:
        hits, cnt = ctx.raycast(scene, rays, any_hit=False, counters=True)
        assert np.array_equal(hits, closest)
        assert np.array_equal(cnt, counters)
        assert np.array_equal(ctx.raycast(scene, rays, any_hit=True), any_hit)
    finally:
        scene.release()
