"""The committed golden fixtures (tests/golden/make_golden.py): the oracle
still produces their hit records bit for bit, the JS host still packs the
same bytes, and the records satisfy what the scenes fix by construction."""
import numpy as np
import pytest

from helpers import load_golden
from oracle.pyoracle import Oracle

FIXTURES = ["cornell_c1.npz", "spheres_c2.npz"]
NONE = 0xFFFFFFFF


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_reproduces_golden_hits(name):
    p, rays, closest, any_hit, counters = load_golden(name)
    o = Oracle(p)
    got, cnt = o.trace(rays, any_hit=False, counters=True)
    assert np.array_equal(got, closest)
    assert np.array_equal(cnt, counters)
    assert np.array_equal(o.trace(rays, any_hit=True), any_hit)


@pytest.mark.parametrize("name", FIXTURES)
def test_golden_hits_are_consistent(name):
    p, rays, closest, any_hit, _ = load_golden(name)
    t = closest[:, 0].view(np.float32)
    hit = closest[:, 1] != NONE
    # Any-hit finds a hit exactly when closest-hit does, never a nearer one.
    assert np.array_equal(any_hit[:, 1] != NONE, hit)
    assert np.all(any_hit[hit, 0].view(np.float32) >= t[hit])
    assert np.all(t[hit] > 1e-4) and np.all(t[hit] < rays[hit, 3])
    assert np.all(closest[hit, 1] < p.primitive_count)
    # Closest hits equal the exhaustive minimum over every primitive.
    brute = Oracle(p).brute_force(rays)
    assert np.array_equal(brute[hit, 0], closest[hit, 0])


def test_cornell_rays_all_hit_the_box():
    """Every primary ray of the closed Cornell box hits something, and the
    centre of the image sees the back wall or an object in front of it."""
    p, rays, closest, _, _ = load_golden("cornell_c1.npz")
    assert np.all(closest[:, 1] != NONE)
    assert len(np.unique(closest[:, 1])) == p.primitive_count


@pytest.mark.parametrize("config,opts,name", [
    ("cornell", {"width": 256, "height": 256}, "cornell_c1.npz"),
    ("spheres", {"scene": {"width": 256, "height": 144}, "width": 256, "height": 144}, "spheres_c2.npz"),
])
def test_host_packs_the_golden_bytes(pack, config, opts, name):
    p = pack(config, **opts)
    g, *_ = load_golden(name)
    for field in ["primitives", "nodes", "materials", "attributes", "lights", "params"]:
        assert np.array_equal(getattr(p, field), getattr(g, field)), field
