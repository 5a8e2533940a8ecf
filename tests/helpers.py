"""Rendering drivers for the tests: the HIP path through the C ABI, and the
oracle, both stepping the per-frame parameters as the Renderer does
(sample offset, running total, shrinking photon radius)."""
from __future__ import annotations

import numpy as np

from raymond_amd import capi
from raymond_amd.scenefile import get_param


def frame_params(p, k: int, params=None) -> np.ndarray:
    """The parameter block for frame k (0-based) of a progressive render."""
    base = p.params if params is None else params
    spp = get_param(base, "samples_per_frame")
    r0 = get_param(p.params, "photon_radius")
    r = r0 * (k + 1) ** (-1.0 / 6.0)
    out = base.copy()
    from raymond_amd.scenefile import set_params
    return set_params(out, sample_offset=k * spp, total_samples=(k + 1) * spp, photon_radius=r, photon_cell=2 * r)


def gpu_render(ctx, p, frames: int = 1, params=None, denoised: bool = False):
    """Renders `frames` frames; returns accum, guide albedo, guide normal (and the denoised image)."""
    scene = ctx.upload_scene(p)
    try:
        ctx.upload_environment(p)
        ctx.resize(p.width, p.height)
        caustics = get_param(p.params if params is None else params, "caustics")
        for k in range(frames):
            ctx.frame(scene, frame_params(p, k, params), p.emitters if caustics else None)
        out = [ctx.read(capi.BUFFER_ACCUM), ctx.read(capi.BUFFER_GUIDE_ALBEDO), ctx.read(capi.BUFFER_GUIDE_NORMAL)]
        if denoised:
            out.append(ctx.read(capi.BUFFER_DENOISED))
        return out
    finally:
        scene.release()


def oracle_render(p, frames: int = 1, params=None, light=True, rows=None, paths=None):
    """The oracle's accumulation after `frames` frames (light tracing and
    photon gather included when the parameters turn caustics on). rows =
    (y0, y1) runs the camera paths of those rows only (light tracing and the
    gather still cover the image, since splats land anywhere): compare that
    band. paths: a light-path budget (RM_OPTION_LIGHT_PATHS) rather than the
    reference's count."""
    from oracle.pyoracle import PHOTON_CAPACITY, PHOTON_TABLE, Oracle
    o = Oracle(p)
    shape = (p.height, p.width, 4)
    acc, ga, gn = (np.zeros(shape, np.float32) for _ in range(3))
    caustics = get_param(p.params if params is None else params, "caustics")
    splats = np.zeros(p.width * p.height * 4, np.uint32)
    for k in range(frames):
        fp = frame_params(p, k, params)
        o.main(fp, acc, ga, gn, rows=rows)
        if caustics and light:
            heads = np.zeros(PHOTON_TABLE + 1, np.uint32)
            if paths is None:
                photons = np.zeros(PHOTON_CAPACITY * 12, np.float32)
                count = min(p.width * p.height * get_param(fp, "samples_per_frame"), PHOTON_CAPACITY)
                o.light_trace(fp, p.emitters, 0, count, splats, heads, photons)
            else:
                photons = np.zeros(paths * 12, np.float32)
                o.light_trace(fp, p.emitters, 0, paths, splats, heads, photons, paths=paths)
            o.resolve_splats(fp, splats, acc)
            o.photon_gather(fp, heads, photons, acc)
    return acc, ga, gn


def agreement(a, b, rel=1e-4, abs_=1e-5) -> float:
    """Fraction of pixels whose four channels all agree within abs + rel·|b|."""
    close = np.abs(a - b) <= abs_ + rel * np.abs(b)
    return float(close.all(axis=-1).mean())


def rel_mse(a, r) -> float:
    """The reference's relMSE: mean over channels of (a - r)² / (r² + 1e-2) (tools/converge.mjs:51-61)."""
    a = a[..., :3].astype(np.float64)
    r = r[..., :3].astype(np.float64)
    return float(np.mean((a - r) ** 2 / (r * r + 1e-2)))


GOLDEN = __import__("pathlib").Path(__file__).resolve().parent / "golden"


def load_golden(name: str):
    """A committed fixture: (Packed scene, rays, expected closest, any-hit, counters).
    The environment is replaced by a 1 × 1 black texel: traversal never reads it."""
    import json
    from raymond_amd.scenefile import Packed
    z = np.load(GOLDEN / name)
    meta = json.loads(bytes(z["meta"]).decode())
    meta.update(envWidth=1, envHeight=1, envTableWidth=1, envTableHeight=0)
    p = Packed(meta=meta, primitives=z["primitives"], nodes=z["nodes"], materials=z["materials"],
               attributes=z["attributes"], lights=z["lights"], params=z["params"], emitters=None,
               env_rgba=np.zeros(4, np.uint16), env_table=np.zeros(1, np.float32))
    return p, z["rays"], z["closest"], z["any_hit"], z["counters"]
