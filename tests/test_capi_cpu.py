"""The C-ABI library and the addon load, and export every entry point
include/raymond.h declares (no GPU calls)."""
import ctypes
import re
import subprocess
from pathlib import Path

import pytest

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
    # More rays than the static scratch holds: refused before anything else.
    assert lib.rm_raycast_static(None, None, None, 65537, 0, None, None) == -1
    assert b"at most 65536 rays" in lib.rm_last_error()
    assert lib.rm_raycast_static(None, None, None, 1, 0, None, None) == -1


def test_addon_exports():
    out = subprocess.run(["node", "-e", "const a=require(process.argv[1]); console.log(Object.keys(a).sort().join(','))",
                          str(ROOT / "raymond_amd" / "lib" / "raymond.node")], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip().split(",") == sorted(
        ["open", "close", "uploadScene", "releaseScene", "uploadTextures", "uploadEnvironment", "resize", "frame",
         "read", "synchronize", "display", "raycast", "deviceInfo", "bvhBuild", "setOption", "serverSetOption",
         "tune", "serverOpen", "serverContext", "serverSubmit", "serverWait", "serverClose", "serverCancel"])


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
    assert lib.rm_server_cancel(None, 1) == -1


def test_bvh_build_on_the_host():
    """rm_bvh_build needs no device: a known two-cluster split, the empty
    tree, and its argument checks."""
    import numpy as np
    from raymond_amd import capi
    lo = np.array([[0, 0, 0], [0.1, 0, 0], [10, 0, 0], [10.1, 0, 0], [10.2, 0, 0]], np.float64)
    nodes, order = capi.bvh_build(lo, lo + 0.05, max_leaf_size=1)
    f = nodes.view(np.float32)
    assert nodes[0, 7] == 0 and nodes[0, 3] == 1, "the root splits into nodes 1 and 2"
    left, right = sorted([1, 2], key=lambda k: f[k, 0])
    assert f[left, 4] < 1 and f[right, 0] >= 10, "the clusters are apart"
    assert sorted(order) == list(range(5))
    assert np.allclose(f[0, :3], 0) and np.allclose(f[0, 4:7], [10.25, 0.05, 0.05])
    nodes, order = capi.bvh_build(np.zeros((0, 3)), np.zeros((0, 3)))
    assert nodes.shape == (1, 8) and not nodes.any() and order.size == 0
    lib = capi.load()
    used = ctypes.c_uint32()
    assert lib.rm_bvh_build(None, None, 1, 4, 0, None, 1, ctypes.byref(used), None) == -1
    buf = (ctypes.c_uint32 * 8)()
    b = (ctypes.c_double * 6)()
    assert lib.rm_bvh_build(b, b, 2, 4, 0, buf, 1, ctypes.byref(used), (ctypes.c_uint32 * 2)()) == -1
    assert b"2" in lib.rm_last_error()


def test_new_entry_points_check_their_arguments():
    """The autotuner, server options and wide quantizer refuse bad arguments
    without touching a device."""
    from raymond_amd import capi
    lib = capi.load()
    assert lib.rm_tune(None, None, None, None, 0, 2, None) == -1
    assert b"rm_tune" in lib.rm_last_error()
    assert lib.rm_server_set_option(None, 0, 0) == -1
    used = ctypes.c_uint32()
    assert lib.rm_bvh_quantize_wide(None, 0, None, 0, ctypes.byref(used)) == -1
    assert lib.rm_set_option(None, capi.OPTION_AUTOTUNE, 1) == -1


def test_python_and_js_constants_match_the_header(tmp_path):
    """capi.py's and addon.mjs's option, schedule, precision, buffer and stage
    numbers are the header's, and the ctypes tuner result has the C layout
    (sizes and offsets printed by a C program built against raymond.h)."""
    import re
    import shutil
    from raymond_amd import capi
    root = Path(__file__).resolve().parent.parent
    header = (root / "include" / "raymond.h").read_text()
    defs = {m.group(1): int(m.group(2)) for m in re.finditer(r"#define (RM_\w+) \(?(-?\d+)\)?", header)}
    py = {f"RM_OPTION_{n}": getattr(capi, f"OPTION_{n}")
          for n in ("SCHEDULE", "PERSISTENT_BLOCKS", "PRECISION", "AUTOTUNE", "BATCH", "WAVES", "LIGHT_PATHS",
                    "STACK_DEPTH", "BLOCK_SHAPE", "SPLIT_DEPTH", "TUNE_CANDIDATES")}
    py.update({f"RM_SCHEDULE_{n}": getattr(capi, f"SCHEDULE_{n}") for n in ("PIXEL", "COMPACT", "PERSISTENT",
                                                                            "WAVEFRONT", "SORTED", "SPLIT")})
    py.update({f"RM_BLOCK_{n}": getattr(capi, f"BLOCK_{n}") for n in ("SQUARES", "ROWS")})
    py["RM_ERR_CANCELLED"] = capi.ERR_CANCELLED
    py.update({f"RM_PRECISION_{n}": getattr(capi, f"PRECISION_{n}") for n in ("QUANTIZED_BVH", "FP16_GUIDES",
                                                                              "COMPENSATED")})
    py.update({f"RM_BUFFER_{n}": getattr(capi, f"BUFFER_{n}") for n in ("ACCUM", "GUIDE_ALBEDO", "GUIDE_NORMAL",
                                                                        "DENOISED")})
    py.update({f"RM_STAGE_{s.upper()}": k for k, s in enumerate(capi.STAGES)})
    py["RM_STAGE_COUNT"] = len(capi.STAGES)
    py["RM_TUNE_CANDIDATES"] = capi.TUNE_CANDIDATES
    assert {k: defs[k] for k in py} == py
    assert max(v for k, v in defs.items() if k.startswith("RM_OPTION_") and k != "RM_OPTION_COUNT") + 1 == \
        defs["RM_OPTION_COUNT"]
    js = (root / "raymond_amd" / "js" / "addon.mjs").read_text()
    for name in ("OPTION_PRECISION", "OPTION_SCHEDULE", "OPTION_WAVES", "OPTION_LIGHT_PATHS"):
        assert int(re.search(rf"export const {name} = (\d+);", js).group(1)) == defs["RM_" + name]
    js_schedules = re.search(r"export const SCHEDULES = Object.freeze\(\{(.*?)\}\);", js).group(1)
    named = {k.strip().strip('"'): int(v) for k, v in (e.split(":") for e in js_schedules.split(","))}
    assert named == capi.SCHEDULES
    if not shutil.which("gcc"):
        pytest.skip("no C compiler")
    src = tmp_path / "layout.c"
    structs = {"rm_tune_result": capi._TuneResult, "rm_job": capi._Job, "rm_job_stats": capi._JobStats}
    body = "".join(f'printf("{c} %zu\\n", sizeof({c}));' +
                   "".join(f'printf("{c}.{f} %zu\\n", offsetof({c}, {f}));' for f, *_ in t._fields_)
                   for c, t in structs.items())
    src.write_text("#include <stddef.h>\n#include <stdio.h>\n#include <raymond.h>\n"
                   f"int main(void) {{ {body} return 0; }}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", f"-I{root / 'include'}", str(src), "-o", str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                       check=True).stdout.splitlines())
    want = {}
    for c, t in structs.items():
        want[c] = str(ctypes.sizeof(t))
        want.update({f"{c}.{f}": str(getattr(t, f).offset) for f, *_ in t._fields_})
    assert got == want
