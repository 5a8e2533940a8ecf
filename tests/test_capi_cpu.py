"""The C-ABI library and the addon load, and export every entry point
include/raymond.h declares (no GPU calls)."""
import ctypes
import re
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def declared():
    text = (ROOT / "include" / "raymond.h").read_text()
    return sorted(set(re.findall(r"^(?:int|const char\*|rm_context\*)\s+(rm_\w+)\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    from raymond_amd import capi
    lib = capi.load()
    names = declared()
    assert len(names) >= 16
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(capi.SIGNATURES), "the ctypes binding covers exactly the header"
    nm = subprocess.run(["nm", "-D", "--defined-only", str(capi.LIB_PATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (rm_\w+)", nm))
    assert exported == set(names), exported ^ set(names)


def test_last_error_starts_empty():
    from raymond_amd import capi
    assert capi.load().rm_last_error() == b""


def test_null_arguments_are_rejected_without_touching_a_device():
    from raymond_amd import capi
    lib = capi.load()
    assert lib.rm_open(0, None) == -1
    assert b"context" in lib.rm_last_error()
    assert lib.rm_close(None) == -1
    assert lib.rm_frame(None, None, None, None, 0) == -1


def test_addon_exports():
    out = subprocess.run(["node", "-e", "const a=require(process.argv[1]); console.log(Object.keys(a).sort().join(','))",
                          str(ROOT / "raymond_amd" / "lib" / "raymond.node")], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip().split(",") == sorted(
        ["open", "close", "uploadScene", "releaseScene", "uploadTextures", "uploadEnvironment", "resize", "frame",
         "read", "synchronize", "display", "raycast", "deviceInfo", "serverOpen", "serverContext", "serverSubmit", "serverWait",
         "serverClose"])


def test_oracle_is_not_linked_into_the_product():
    lib = ROOT / "raymond_amd" / "lib" / "libraymond_hip.so"
    nm = subprocess.run(["nm", "-D", str(lib)], capture_output=True, text=True).stdout
    assert "oracle_" not in nm
    ldd = subprocess.run(["ldd", str(lib)], capture_output=True, text=True).stdout
    assert "liboracle" not in ldd


def test_server_rejects_null_arguments_without_touching_a_device():
    from raymond_amd import capi
    lib = capi.load()
    assert lib.rm_server_open(0, 4, 1 << 30, None) == -1
    assert b"server" in lib.rm_last_error()
    assert lib.rm_server_context(None) is None
    assert lib.rm_server_submit(None, None, None) == -1
    assert lib.rm_server_wait(None, 1, None, 0, None) == -1
    assert lib.rm_server_close(None) == -1
