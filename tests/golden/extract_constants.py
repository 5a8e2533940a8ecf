"""Writes tests/golden/shader_constants.json: the numeric literals of the
reference's WGSL (src/shaders.mjs), read as text — every top-level `const`
and, for every `fn`, the literals in its body in order of appearance.

    python tests/golden/extract_constants.py [/root/reference/src/shaders.mjs]

The fixture is data (names and numbers); tests/test_constants_cpu.py checks
the oracle and the kernels against it, so a digit dropped from a hash
multiplier, a threshold or a filter weight in either fails there.
"""
from __future__ import annotations

import json
import re
import sys
from pathlib import Path

OUT = Path(__file__).resolve().parent / "shader_constants.json"
SOURCE = Path("/root/reference/src/shaders.mjs")

# A WGSL numeric literal: hex or decimal, optional exponent, optional suffix.
NUMBER = re.compile(r"(?<![\w.])(0[xX][0-9a-fA-F]+[iu]?|(?:\d+\.\d*|\.\d+|\d+)(?:[eE][+-]?\d+)?[iuf]?)(?![\w.])")


def literal_value(text: str):
    """(value, kind) of a literal: kind 'u' / 'i' for integers, 'f' for floats."""
    if text.lower().startswith("0x"):
        t = text.rstrip("iu")
        return int(t, 16), text[len(t):] or "i"
    t = text.rstrip("iuf")
    suffix = text[len(t):]
    if suffix in ("u", "i") or re.fullmatch(r"\d+", t):
        return int(t), suffix or "i"
    return float(t), "f"


def strip_comments(src: str) -> str:
    return re.sub(r"//[^\n]*", "", src)


def function_bodies(src: str) -> dict:
    """{name: body text} for every `fn name(...) ... { body }` in WGSL source."""
    out = {}
    for m in re.finditer(r"\bfn\s+(\w+)\s*\(", src):
        start = src.index("{", m.end())
        depth, i = 0, start
        while True:
            c = src[i]
            depth += c == "{"
            depth -= c == "}"
            i += 1
            if depth == 0:
                break
        out[m.group(1)] = src[start:i]
    return out


def extract(path: Path) -> dict:
    text = path.read_text()
    # The WGSL lives in template strings; the JS around them has no `fn`.
    wgsl = strip_comments(text)
    consts = {}
    for m in re.finditer(r"^\s*const\s+(\w+)\s*(?::\s*\w+)?\s*=\s*([^;`]+);", wgsl, re.M):
        name, value = m.group(1), m.group(2).strip()
        nums = NUMBER.findall(value)
        if nums and not value.startswith(("true", "false", "/*")):
            consts[name] = [literal_value(n) for n in nums]
    functions = {}
    for name, body in function_bodies(wgsl).items():
        functions.setdefault(name, [literal_value(n) for n in NUMBER.findall(body)])
    return {"source": "src/shaders.mjs", "consts": consts, "functions": functions}


if __name__ == "__main__":
    src = Path(sys.argv[1]) if len(sys.argv) > 1 else SOURCE
    data = extract(src)
    OUT.write_text(json.dumps(data, indent=1, sort_keys=True) + "\n")
    print(f"wrote {OUT}: {len(data['consts'])} constants, {len(data['functions'])} functions")
