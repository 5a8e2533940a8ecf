"""Builds the native parts in-tree:

* ``raymond_amd/lib/libraymond_hip.so`` — the HIP kernels and C ABI, for gfx950;
* ``raymond_amd/lib/raymond.node`` — the N-API addon the JS Renderer loads;
* ``oracle/build/liboracle.so`` — the CPU oracle (test infrastructure only).

Run ``python -m raymond_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "raymond_amd"
CSRC = PKG / "csrc"
LIB = PKG / "lib"
OBJ = PKG / "_build"
ORACLE = ROOT / "oracle"

ARCH = os.environ.get("RAYMOND_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

# FP32 exactly as written: no FMA contraction, IEEE division and square root,
# and no packed-FP32 instructions. No SLP vectorisation either: with it, the
# persistent path kernel's guide normal at light hits came out as the value
# from before the assignment, although the optimised IR selects the new one
# (a <4 x float> phi across a divergent switch, lowered wrongly for gfx950);
# without it the kernel matches the one-thread-per-pixel kernel bit for bit.
# The fault is not confined to that kernel: built with SLP, main.hip's
# k_main for volumes gives other guide normals than the oracle on 1.8% of the
# fog-and-lamp fixture's pixels. So it stays off for every file.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-fno-slp-vectorize",
    f"-I{ROOT / 'include'}", "-Wno-unused-command-line-argument",
]
SOURCES = ["main.hip", "main5.hip", "main_shapes.hip", "split.hip", "split5.hip", "batch.hip", "schedules.hip",
           "light.hip", "kernels.hip", "api.hip", "server.hip"]
# Host-only C++ (the BVH builder): double precision as written, no FMA contraction.
CXX_SOURCES = ["bvh_build.cpp"]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-pthread", "-ffp-contract=off", "-fno-fast-math", "-Wall",
             f"-I{ROOT / 'include'}"]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")


def build_hip() -> Path:
    OBJ.mkdir(parents=True, exist_ok=True)
    LIB.mkdir(parents=True, exist_ok=True)
    objs = [OBJ / (s.replace(".hip", ".o")) for s in SOURCES]
    cmds = [[HIPCC, *HIP_FLAGS, "-c", str(CSRC / s), "-o", str(o)] for s, o in zip(SOURCES, objs)]
    cxx_objs = [OBJ / (s.replace(".cpp", ".o")) for s in CXX_SOURCES]
    cmds += [["g++", *CXX_FLAGS, "-c", str(CSRC / s), "-o", str(o)] for s, o in zip(CXX_SOURCES, cxx_objs)]
    with ThreadPoolExecutor(max_workers=len(cmds)) as pool:
        list(pool.map(_run, cmds))
    out = LIB / "libraymond_hip.so"
    tmp = LIB / "libraymond_hip.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", *map(str, objs + cxx_objs), "-o",
          str(tmp)])
    tmp.replace(out)
    return out


def build_addon() -> Path:
    out = LIB / "raymond.node"
    tmp = LIB / "raymond.node.tmp"
    _run(["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-Wall", "-I/usr/include/node", f"-I{ROOT / 'include'}",
          str(PKG / "addon" / "raymond_napi.c"), "-o", str(tmp), f"-L{LIB}", "-lraymond_hip",
          "-Wl,-rpath,$ORIGIN"])
    tmp.replace(out)
    return out


def build_oracle() -> Path:
    (ORACLE / "build").mkdir(parents=True, exist_ok=True)
    out = ORACLE / "build" / "liboracle.so"
    tmp = ORACLE / "build" / "liboracle.so.tmp"
    _run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-shared", "-fPIC", "-pthread",
          str(ORACLE / "oracle.cpp"), "-o", str(tmp)])
    tmp.replace(out)
    return out


def build_all(verbose: bool = True) -> None:
    with ThreadPoolExecutor(max_workers=2) as pool:
        hip = pool.submit(build_hip)
        oracle = pool.submit(build_oracle)
        hip.result()
        oracle.result()
    build_addon()
    if verbose:
        for p in (LIB / "libraymond_hip.so", LIB / "raymond.node", ORACLE / "build" / "liboracle.so"):
            print(f"built {p.relative_to(ROOT)} ({p.stat().st_size // 1024} KiB)")


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    {"all": build_all, "hip": build_hip, "addon": build_addon, "oracle": build_oracle}[what]()
