"""Workload scenes for Python harnesses.

Scenes are described, packed and BVH-built by the JS host modules (the same
code the Renderer uploads from); ``raymond_amd/js/tools/pack.mjs`` writes
the buffers to one ``.rmpk`` file, which this module reads back.
"""
from __future__ import annotations

import hashlib
import json
import os
import struct
import subprocess
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
PACK_TOOL = ROOT / "raymond_amd" / "js" / "tools" / "pack.mjs"
CACHE = Path(os.environ.get("RAYMOND_SCENE_CACHE", ROOT / "_build" / "scenes"))
_DTYPES = {"u32": np.uint32, "f32": np.float32, "u8": np.uint8, "u16": np.uint16}


@dataclass
class Packed:
    meta: dict
    primitives: np.ndarray          # u32, 16 per slot
    nodes: np.ndarray               # u32, 8 per node
    materials: np.ndarray           # u32, 24 per material
    attributes: np.ndarray          # f32
    lights: np.ndarray              # u32 table
    params: np.ndarray              # 208 bytes
    emitters: np.ndarray | None     # f32 or None
    env_rgba: np.ndarray            # u16
    env_table: np.ndarray           # f32
    textures: tuple | None = None       # (RGBA8 layers, width, height, layers), sRGB colour
    normal_maps: tuple | None = None    # (RGBA8 layers, width, height, layers), plain
    extra: dict = field(default_factory=dict)

    @property
    def primitive_count(self) -> int:
        return int(self.meta["primitiveCount"])

    @property
    def width(self) -> int:
        return int(self.meta["width"])

    @property
    def height(self) -> int:
        return int(self.meta["height"])

    def with_params(self, **fields) -> np.ndarray:
        """A copy of the parameter block with named fields changed."""
        return set_params(self.params.copy(), **fields)


# Word offsets of the parameter block's scalar fields (src/shaders.mjs:9-24).
PARAM_FIELDS = {
    "lens_radius": (3, "f"), "tan_half_fov": (7, "f"), "focus_distance": (11, "f"), "aspect": (15, "f"),
    "sun_omc": (19, "f"), "sun_on": (23, "u"), "firefly_clamp": (27, "f"), "light_count": (31, "u"),
    "max_bounces": (35, "u"), "view": (36, "u"), "volume_count": (37, "u"), "env_strength": (38, "f"),
    "env_rotation": (39, "f"), "width": (40, "u"), "height": (41, "u"), "sample_offset": (42, "u"),
    "samples_per_frame": (43, "u"), "total_samples": (44, "u"), "exposure": (45, "f"), "denoise": (46, "u"),
    "caustics": (47, "u"), "photon_radius": (48, "f"), "photon_cell": (49, "f"),
}


def set_params(params: np.ndarray, **fields) -> np.ndarray:
    for name, value in fields.items():
        word, kind = PARAM_FIELDS[name]
        if kind == "u":
            params[word * 4:word * 4 + 4] = np.frombuffer(struct.pack("<I", int(value)), np.uint8)
        else:
            params[word * 4:word * 4 + 4] = np.frombuffer(struct.pack("<f", float(value)), np.uint8)
    return params


def resized_params(params: np.ndarray, width: int, height: int, max_bounces: int | None = None) -> np.ndarray:
    """The parameter block for the same scene and camera at another image
    size: size, aspect and the starting photon radius (about three pixels
    across at the delta objects, so inversely proportional to the height;
    raymond_amd/js/frame.mjs writeParams / writeFrameParams), and optionally
    the bounce limit."""
    out = params.copy()
    h0 = get_param(params, "height")
    r = max(float(get_param(params, "photon_radius")) * h0 / height, 1e-5)
    fields = dict(width=width, height=height, aspect=width / height, photon_radius=r, photon_cell=2 * r)
    if max_bounces is not None:
        fields["max_bounces"] = max_bounces
    return set_params(out, **fields)


def get_param(params: np.ndarray, name: str):
    word, kind = PARAM_FIELDS[name]
    return struct.unpack("<I" if kind == "u" else "<f", params[word * 4:word * 4 + 4].tobytes())[0]


def read(path: str | Path) -> Packed:
    data = Path(path).read_bytes()
    if data[:8] != b"RMPACK01":
        raise ValueError(f"{path} is not a packed Raymond scene")
    (n,) = struct.unpack_from("<I", data, 8)
    meta = json.loads(data[12:12 + n].decode())
    base = ((12 + n + 15) // 16) * 16
    arrays = {}
    for s in meta["sections"]:
        raw = data[base + s["offset"]: base + s["offset"] + s["bytes"]]
        arrays[s["name"]] = np.frombuffer(raw, dtype=_DTYPES[s["dtype"]]).copy()
    tex = (arrays["textures"], *meta["textures"]) if meta.get("textures") else None
    nmap = (arrays["normal_maps"], *meta["normalMaps"]) if meta.get("normalMaps") else None
    return Packed(meta=meta, primitives=arrays["primitives"], nodes=arrays["nodes"], materials=arrays["materials"],
                  attributes=arrays["attributes"], lights=arrays["lights"], params=arrays["params"],
                  emitters=arrays.get("emitters"), env_rgba=arrays["env_rgba"], env_table=arrays["env_table"],
                  textures=tex, normal_maps=nmap, extra={"object_of": arrays.get("object_of")})


def pack(config: str, options: dict | None = None, cache: Path | None = None, timeout: float = 600) -> Packed:
    """Packs a named workload (see raymond_amd/js/configs.mjs) with node, cached by its options."""
    options = options or {}
    key = hashlib.sha1(json.dumps([config, options], sort_keys=True).encode()).hexdigest()[:16]
    sources = sorted((ROOT / "raymond_amd" / "js").rglob("*.mjs"))
    stamp = hashlib.sha1(b"".join(p.read_bytes() for p in sources)).hexdigest()[:8]
    cache = cache or CACHE
    cache.mkdir(parents=True, exist_ok=True)
    out = cache / f"{config}-{key}-{stamp}.rmpk"
    if not out.exists():
        tmp = out.with_suffix(".tmp")
        res = subprocess.run(["node", str(PACK_TOOL), config, str(tmp), json.dumps(options)],
                             capture_output=True, text=True, timeout=timeout)
        if res.returncode != 0:
            raise RuntimeError(f"packing {config} failed:\n{res.stderr}")
        tmp.replace(out)
    return read(out)
