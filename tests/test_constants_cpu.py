"""The reference shader's numeric constants in the oracle and in the kernels.

tests/golden/shader_constants.json holds every numeric literal of the
reference's WGSL (src/shaders.mjs), function by function, as
tests/golden/extract_constants.py read them from the source text. Here:

* each reference function's literals appear in the oracle's restatement of
  it and in the kernels' (a hash multiplier, threshold or filter weight
  with a digit wrong fails), apart from the listed exceptions, each with
  its reason and, where a literal was folded into a named constant, a
  check of that constant;
* the PCG hash, the Owen scramble and the photon-grid hash, built in numpy
  from the fixture's constants (the algorithms' published structure:
  Jarzynski & Olano 2020's PCG hash, Burley 2020's hash-based Owen
  scrambling with Laine & Karras's mixing steps), equal the oracle's on
  every input tried.
"""
import json
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import pyoracle

ROOT = Path(__file__).resolve().parent.parent
FIXTURE = json.loads((ROOT / "tests" / "golden" / "shader_constants.json").read_text())
ORACLE = ROOT / "oracle" / "oracle.cpp"
KERNELS = [ROOT / "raymond_amd" / "csrc" / f for f in ("device.hpp", "shading.hpp", "path.hpp", "light.hip",
                                                        "kernels.hip", "kernels.hpp")]

CNUM = re.compile(r"(?<![\w.])(0[xX][0-9a-fA-F]+[uUlL]*|(?:\d+\.\d*|\.\d+|\d+)(?:[eE][+-]?\d+)?[fFuUlL]*)(?![\w.])")


def c_value(t):
    if t.lower().startswith("0x"):
        return int(t.rstrip("uUlL"), 16)
    s = t.rstrip("fFuUlL")
    if re.fullmatch(r"\d+", s) and not t.lower().endswith("f"):
        return int(s)
    return float(s)


def bodies(paths):
    """{name: [bodies]} of the C++ functions defined in the files."""
    text = re.sub(r"//[^\n]*", "", "".join(Path(p).read_text() for p in paths))
    out = {}
    for m in re.finditer(r"\b(\w+)\s*\(([^;{}()]|\([^)]*\))*\)\s*(const\s*)?\{", text):
        name = m.group(1)
        if name in ("if", "for", "while", "switch", "return"):
            continue
        i = m.end() - 1
        depth = 0
        start = i
        while True:
            depth += text[i] == "{"
            depth -= text[i] == "}"
            i += 1
            if depth == 0:
                break
        out.setdefault(name, []).append(text[start:i])
    return out


def same(a, b):
    if isinstance(a, int) and isinstance(b, int):
        return a == b
    return np.float32(a) == np.float32(b)


# The reference function -> where the oracle and the kernels restate it,
# when the name differs.
ORACLE_NAMES = {"aces": ["aces1"], "to_srgb": ["srgb1"], "atrous": ["oracle_atrous"], "prepare": ["oracle_prepare"],
                "compose": ["oracle_compose"], "main": ["main_pixel"], "light_trace": ["light_trace_one"],
                "photon_gather": ["photon_gather_pixel"], "resolve_splats": ["oracle_resolve_splats"],
                "qmc_next": ["next"]}
KERNEL_NAMES = {"aces": ["aces1"], "to_srgb": ["srgb1"], "atrous": ["atrous_pixel"], "prepare": ["k_prepare"],
                "compose": ["compose_pixel"], "main": ["main_pixel", "path_begin"], "light_trace": ["light_paths"],
                "photon_gather": ["k_gather_begin", "gather_bounce"], "gather": ["gather_lane", "gather_cell", "photon_term"],
                "resolve_splats": ["resolve_add"], "qmc_next": ["next"],
                "finite": ["finite3"], "radiance": ["path_step", "path_shade", "path_begin"],
                "deposit": ["light_paths"], "dispersed": ["path_shade", "dispersed_ior"],
                "env_choice": ["light_choices", "env_choice"], "sun_choice": ["light_choices", "sun_choice"],
                "occluded": ["occluded", "shadow"]}

# Literals that appear in neither restatement as written, and why.
TWO = {2.0, 2, 4.0}   # 2π, 4π and π/2 are folded once, in double, into TWO_PI, FOUR_PI, PI_2 (as Tint folds them)
EXCEPTIONS = {
    "concentric_disk": TWO, "cosine_hemisphere": TWO, "env_pdf": TWO, "env_uv": TWO, "henyey_greenstein": TWO,
    "light_pdf": TWO, "sample_cone": TWO, "sample_light": TWO, "sphere_uv": TWO, "sample_ggx": TWO | {1.0, 0.0},
    "intersect": TWO,   # the box's slab bounds are ±1 after a subtraction the restatement writes as s3(-1)
    "inspect": TWO,     # pixel centre +0.5 is written as a product with 0.5
    "dispersed": {0.5876},   # 1/0.5876² folded into INV_D_LINE2
    "env_cdf": {0}, "sample_env": {1}, "transmittance": {1.0},
    # The kernels compute x⁵ in Schlick's term by multiplication rather than pow(x, 5).
    "schlick": {0.0, 5.0},
    # Sobol's second dimension in the kernels is a closed form (a superset-XOR
    # transform and a bit reversal), proved equal to the bit walk on the CPU.
    "sobol2": {0, 2147483648},
    # The kernels read a pixel's splats as one uint4 rather than words index·4 + k.
    "resolve_splats": {4, 1, 2},
    # The reference's photons are 3 vec4s each (index · 3); the kernels' sorted
    # photon map keeps positions and (power, direction) pairs in arrays of their own.
    "gather": {3},
    # The stack bound 64 is STACK_LIMIT; the restatements start from named values.
    "trace": {0.0, 0, 1.0, 1, 64},
}
# Reference functions with no counterpart to compare (the vertex stage of
# the display, whose image the kernels write directly; folded helpers).
NOT_RESTATED = {"vertex", "fragment", "inside", "samples", "light_at", "volume", "radiance", "main", "photon_gather",
                "light_trace", "deposit", "env_choice", "sun_choice", "occluded", "dispersed", "finite", "qmc_next"}


def check(paths, names):
    ours = bodies(paths)
    problems = []
    for name, lits in sorted(FIXTURE["functions"].items()):
        found = [b for n in names.get(name, [name]) for b in ours.get(n, [])]
        if not found:
            if name not in NOT_RESTATED:
                problems.append(f"{name}: no restatement found")
            continue
        vals = [c_value(t) for b in found for t in CNUM.findall(b)]
        allowed = EXCEPTIONS.get(name, set())
        missing = [v for v, _ in lits if not any(same(v, w) for w in vals) and v not in allowed]
        if missing:
            problems.append(f"{name}: {missing}")
    return problems


def test_fixture_holds_the_known_constants():
    f = FIXTURE["functions"]
    assert [v for v, _ in f["pcg"]] == [747796405, 2891336453, 28, 4, 277803737, 22]
    assert [v for v, _ in f["photon_cell_hash"]][:3] == [73856093, 19349663, 83492791]
    assert [v for v, _ in FIXTURE["consts"]["KERNEL"]][1:] == [0.375, 0.25, 0.0625]
    assert len(f) > 80


def test_oracle_restates_every_constant():
    problems = check([ORACLE], ORACLE_NAMES)
    assert not problems, problems


def test_kernels_restate_every_constant():
    problems = check(KERNELS, KERNEL_NAMES)
    assert not problems, problems


def test_top_level_constants():
    consts = {k: [v for v, _ in vs] for k, vs in FIXTURE["consts"].items()}
    device = (ROOT / "raymond_amd" / "csrc" / "device.hpp").read_text()
    kernels = (ROOT / "raymond_amd" / "csrc" / "kernels.hip").read_text()
    oracle = ORACLE.read_text()
    for name in ("T_MIN", "INF", "PHOTON_TABLE", "PHOTON_CAPACITY", "SPLAT_SCALE", "NONE"):
        (want,) = consts[name]
        for text, where in ((device, "device.hpp"), (oracle, "oracle.cpp")):
            m = re.search(rf"\b{name}\s*=\s*([^;]+);", text)
            assert m, (name, where)
            got = [c_value(t) for t in CNUM.findall(m.group(1))]
            assert any(same(want, g) for g in got), (name, where, want, got)
    kernel = consts["KERNEL"][1:]
    assert [float(x) for x in re.findall(r"KERNEL[012] = ([0-9.]+)f", kernels)] == kernel
    assert [float(x) for x in re.search(r"KERNEL\[3\] = \{([^}]*)\}", oracle).group(1).replace("f", "").split(",")] \
        == kernel
    # Folded constants, as WGSL folds them: in double, rounded once.
    (pi,) = consts["PI"]
    for text in (device, oracle):
        assert re.search(r"PI_D\s*=\s*3\.14159265358979\b", text)
    assert np.float32(2.0 * pi) == np.float32(6.2831853071795800)
    assert re.search(r"INV_D_LINE2 = \(float\)\(1\.0 / \(0\.5876 \* 0\.5876\)\)", device)
    assert re.search(r"STACK_LIMIT = 64u", device)


M32 = np.uint64(0xFFFFFFFF)


def pcg_numpy(v):
    """Jarzynski & Olano's PCG hash, with the fixture's constants."""
    a, b, s1, s2, c, s3 = [np.uint64(x) for x, _ in FIXTURE["functions"]["pcg"]]
    state = (v * a + b) & M32
    word = (((state >> ((state >> s1) + s2)) ^ state) * c) & M32
    return (word >> s3) ^ word


def reverse_bits(v):
    out = np.zeros_like(v)
    for i in range(32):
        out |= ((v >> np.uint64(i)) & np.uint64(1)) << np.uint64(31 - i)
    return out


def owen_numpy(x, seed):
    """Burley's hash-based Owen scramble (Laine & Karras's steps on the
    reversed bits), with the fixture's constants."""
    k1, shift, one, k2, k3 = [np.uint64(v) for v, _ in FIXTURE["functions"]["owen"]]
    v = reverse_bits(x)
    v ^= (v * k1) & M32
    v = (v + seed) & M32
    v = (v * ((seed >> shift) | one)) & M32
    v ^= (v * k2) & M32
    v ^= (v * k3) & M32
    return reverse_bits(v)


def test_pcg_equals_the_published_hash_with_the_reference_constants():
    rng = np.random.default_rng(11)
    xs = np.concatenate([np.arange(4096, dtype=np.uint64), rng.integers(0, 1 << 32, 20000, dtype=np.uint64),
                         np.array([0xFFFFFFFF, 0x80000000], np.uint64)])
    want = pcg_numpy(xs)
    got = np.array([pyoracle.pcg(int(x)) for x in xs], np.uint64)
    np.testing.assert_array_equal(got, want)


def test_owen_equals_the_published_scramble_with_the_reference_constants():
    rng = np.random.default_rng(12)
    xs = rng.integers(0, 1 << 32, 20000, dtype=np.uint64)
    seeds = rng.integers(0, 1 << 32, 20000, dtype=np.uint64)
    want = owen_numpy(xs, seeds)
    got = np.array([pyoracle.owen(int(x), int(s)) for x, s in zip(xs, seeds)], np.uint64)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("path", [ORACLE, ROOT / "raymond_amd" / "csrc" / "light.hip"], ids=["oracle", "kernels"])
def test_photon_hash_primes(path):
    """The photon grid's cell hash multiplies x, y, z by the fixture's three
    primes and XORs them (Teschner et al. 2003's spatial hash)."""
    primes = [v for v, _ in FIXTURE["functions"]["photon_cell_hash"]][:3]
    (body,) = bodies([path])["photon_cell_hash"]
    assert [int(x) for x in re.findall(r"\*\s*(\d+)u\)", body)] == primes
    assert body.count("^") == 2
