"""Runs the JS host modules' tests (tests/js/host_test.mjs) under node."""
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_js_host_modules():
    res = subprocess.run(["node", str(ROOT / "tests" / "js" / "host_test.mjs")], capture_output=True, text=True,
                         timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "0 failed" in res.stdout


def test_generators_keep_their_pinned_bytes():
    """Texture and Mesh outputs and the packed scenes built from them equal
    the digests in tests/golden/js_generators.json, byte for byte."""
    res = subprocess.run(["node", str(ROOT / "tests" / "golden" / "make_js_fixtures.mjs"), "--check"],
                         capture_output=True, text=True, timeout=600)
    assert res.returncode == 0, res.stdout + res.stderr
