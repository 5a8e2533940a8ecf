"""Runs the JS host modules' tests (tests/js/host_test.mjs) under node."""
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_js_host_modules():
    res = subprocess.run(["node", str(ROOT / "tests" / "js" / "host_test.mjs")], capture_output=True, text=True,
                         timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "0 failed" in res.stdout
