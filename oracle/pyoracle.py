"""ctypes binding of the CPU oracle (oracle/oracle.cpp) — TEST INFRASTRUCTURE ONLY.

Imported by tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline
leg, to check and time the HIP path; never by the product.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "build" / "liboracle.so"
PHOTON_TABLE = 1 << 20
PHOTON_CAPACITY = 1 << 19


class OracleScene(C.Structure):
    _fields_ = [
        ("prims", C.c_void_p), ("prim_slots", C.c_uint32),
        ("nodes", C.c_void_p), ("node_count", C.c_uint32),
        ("materials", C.c_void_p), ("material_count", C.c_uint32),
        ("attributes", C.c_void_p), ("attribute_vec4s", C.c_uint32),
        ("lights", C.c_void_p), ("lights_len", C.c_uint32),
        ("env_rgba", C.c_void_p), ("env_width", C.c_uint32), ("env_height", C.c_uint32),
        ("env_table", C.c_void_p), ("table_width", C.c_uint32), ("table_height", C.c_uint32),
        ("textures", C.c_void_p), ("tex_width", C.c_uint32), ("tex_height", C.c_uint32), ("tex_layers", C.c_uint32),
        ("normal_maps", C.c_void_p), ("nmap_width", C.c_uint32), ("nmap_height", C.c_uint32),
        ("nmap_layers", C.c_uint32),
    ]


_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built: run python -m raymond_amd.build oracle")
        lib = C.CDLL(str(LIB_PATH))
        P, U = C.c_void_p, C.c_uint32
        sig = {
            "oracle_trace": (None, [P, P, U, C.c_int, P, P, C.c_int]),
            "oracle_trace64": (None, [P, P, U, P, C.c_int]),
            "oracle_camera_rays": (None, [P, P]),
            "oracle_brute_force": (None, [P, P, U, U, P]),
            "oracle_main": (None, [P, P, P, P, P, U, U, C.c_int]),
            "oracle_path_stats": (None, [P, P, U, P]),
            "oracle_light_trace": (None, [P, P, P, U, U, P, P, P]),
            "oracle_light_trace_budget": (None, [P, P, P, U, U, U, U, P, P, P]),
            "oracle_resolve_splats": (None, [P, P, P]),
            "oracle_photon_gather": (None, [P, P, P, P, P, C.c_int]),
            "oracle_prepare": (None, [P, P, P, P, P]),
            "oracle_atrous": (None, [P, U, P, P, P]),
            "oracle_compose": (None, [P, P, P, P]),
            "oracle_denoise64": (None, [P, P, P, P, P]),
            "oracle_display": (None, [P, P, P, P]),
            "oracle_pcg": (U, [U]),
            "oracle_owen": (U, [U, U]),
            "oracle_sobol2": (U, [U]),
            "oracle_reverse_bits": (U, [U]),
            "oracle_qmc_draws": (None, [U, U, U, U, C.c_int, P]),
            "oracle_intersect": (None, [P, U, P, C.c_float, P]),
            "oracle_eval_bsdf": (None, [P, P, P, P, P, P]),
            "oracle_fresnel_dielectric": (C.c_float, [C.c_float, C.c_float]),
            "oracle_ggx_terms": (None, [C.c_float, C.c_float, C.c_float, C.c_float, P]),
            "oracle_ggx_sample": (None, [C.c_float, C.c_float, C.c_float, P]),
            "oracle_sample_bsdf": (None, [P, P, P, P, U, P]),
            "oracle_wavelength_rgb": (None, [C.c_float, P]),
            "oracle_env_pdf": (C.c_float, [P, P, P]),
            "oracle_env_radiance": (None, [P, P, P, P]),
        }
        for name, (res, argt) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = argt
        _lib = lib
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Oracle:
    """The oracle over one packed scene (a raymond_amd.scenefile.Packed)."""

    def __init__(self, packed, textures=None, normal_maps=None):
        self.packed = packed
        textures = textures if textures is not None else getattr(packed, "textures", None)
        normal_maps = normal_maps if normal_maps is not None else getattr(packed, "normal_maps", None)
        m = packed.meta
        self._keep = [packed.primitives, packed.nodes, packed.materials, packed.attributes, packed.lights,
                      packed.env_rgba, packed.env_table]
        s = OracleScene()
        s.prims, s.prim_slots = _p(packed.primitives), packed.primitives.size // 16
        s.nodes, s.node_count = _p(packed.nodes), packed.nodes.size // 8
        s.materials, s.material_count = _p(packed.materials), packed.materials.size // 24
        s.attributes, s.attribute_vec4s = _p(packed.attributes), packed.attributes.size // 4
        s.lights, s.lights_len = _p(packed.lights), packed.lights.size
        s.env_rgba, s.env_width, s.env_height = _p(packed.env_rgba), m["envWidth"], m["envHeight"]
        s.env_table, s.table_width, s.table_height = _p(packed.env_table), m["envTableWidth"], m["envTableHeight"]
        if textures is not None:
            data, w, h, n = textures
            self._keep.append(data)
            s.textures, s.tex_width, s.tex_height, s.tex_layers = _p(data), w, h, n
        if normal_maps is not None:
            data, w, h, n = normal_maps
            self._keep.append(data)
            s.normal_maps, s.nmap_width, s.nmap_height, s.nmap_layers = _p(data), w, h, n
        self.s = s
        self.lib = load()

    @property
    def ref(self):
        return C.byref(self.s)

    def trace(self, rays: np.ndarray, any_hit=False, counters=False, threads=8):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        hits = np.empty((r.shape[0], 4), np.uint32)
        cnt = np.empty((r.shape[0], 2), np.uint32) if counters else None
        self.lib.oracle_trace(self.ref, _p(r), r.shape[0], int(any_hit), _p(hits), _p(cnt), threads)
        return (hits, cnt) if counters else hits

    def trace64(self, rays: np.ndarray, threads=8):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        hits = np.empty((r.shape[0], 4), np.float64)
        self.lib.oracle_trace64(self.ref, _p(r), r.shape[0], _p(hits), threads)
        return hits

    def brute_force(self, rays: np.ndarray):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        out = np.empty((r.shape[0], 2), np.uint32)
        self.lib.oracle_brute_force(self.ref, _p(r), r.shape[0], self.packed.primitive_count, _p(out))
        return out

    def camera_rays(self, params=None):
        params = self.packed.params if params is None else params
        w, h = self.packed.width, self.packed.height
        rays = np.empty((h * w, 8), np.float32)
        self.lib.oracle_camera_rays(_p(np.ascontiguousarray(params)), _p(rays))
        return rays

    def main(self, params, accum, galb, gnorm, rows=None, threads=8):
        h = int(np.frombuffer(params[164:168].tobytes(), np.uint32)[0])
        y0, y1 = rows if rows else (0, h)
        self.lib.oracle_main(self.ref, _p(params), _p(accum), _p(galb), _p(gnorm), y0, y1, threads)

    def path_stats(self, params, stride=1):
        counts = np.zeros(7, np.uint64)
        self.lib.oracle_path_stats(self.ref, _p(params), stride, _p(counts))
        keys = ("closest", "shadow", "nodes", "prims", "surfaces", "smooth", "samples")
        return dict(zip(keys, (int(c) for c in counts)))

    def light_trace(self, params, emitters, first, count, splats, heads, photons, paths=None, capacity=None):
        """Light paths [first, first + count) of a frame of `paths` (default:
        the reference's count), depositing into room for `capacity` photons."""
        if paths is None:
            self.lib.oracle_light_trace(self.ref, _p(params), _p(emitters), first, count, _p(splats), _p(heads),
                                        _p(photons))
        else:
            self.lib.oracle_light_trace_budget(self.ref, _p(params), _p(emitters), first, count, paths,
                                               paths if capacity is None else capacity, _p(splats), _p(heads),
                                               _p(photons))

    def resolve_splats(self, params, splats, accum):
        self.lib.oracle_resolve_splats(_p(params), _p(splats), _p(accum))

    def photon_gather(self, params, heads, photons, accum, threads=8):
        self.lib.oracle_photon_gather(self.ref, _p(params), _p(heads), _p(photons), _p(accum), threads)

    def denoise(self, params, accum, galb, gnorm):
        """prepare, à-trous at steps 1..16, compose: returns every pass's output."""
        lib = self.lib
        a = np.zeros_like(accum)
        b = np.zeros_like(accum)
        lib.oracle_prepare(_p(params), _p(accum), _p(galb), _p(gnorm), _p(a))
        passes = [a.copy()]
        src, dst = a, b
        for step in (1, 2, 4, 8, 16):
            lib.oracle_atrous(_p(params), step, _p(gnorm), _p(src), _p(dst))
            passes.append(dst.copy())
            src, dst = dst, src
        out = np.zeros_like(accum)
        lib.oracle_compose(_p(params), _p(galb), _p(src), _p(out))
        passes.append(out)
        return passes

    @staticmethod
    def denoise64(params, accum, galb, gnorm):
        """The denoiser in FP64 from the FP32 sums: (H, W, 4) float64."""
        out = np.zeros(accum.shape, np.float64)
        load().oracle_denoise64(_p(np.ascontiguousarray(params)), _p(np.ascontiguousarray(accum, np.float32)),
                                _p(np.ascontiguousarray(galb, np.float32)), _p(np.ascontiguousarray(gnorm, np.float32)),
                                _p(out))
        return out

    def intersect(self, prim, ray, t_max=1e30):
        out = np.zeros(3, np.float32)
        r = np.ascontiguousarray(ray, np.float32)
        self.lib.oracle_intersect(self.ref, prim, _p(r), t_max, _p(out))
        return out


def display(params: np.ndarray, accum: np.ndarray, denoised: np.ndarray | None = None) -> np.ndarray:
    """DISPLAY_SHADER as RGBA8 (oracle_display) for an accumulation and, when
    the parameters denoise, the denoised average."""
    p = np.ascontiguousarray(params, dtype=np.uint8)
    a = np.ascontiguousarray(accum, dtype=np.float32)
    d = np.ascontiguousarray(a if denoised is None else denoised, dtype=np.float32)
    out = np.empty(a.shape[:-1] + (4,), np.uint8)
    load().oracle_display(_p(p), _p(a), _p(d), _p(out))
    return out


def pcg(v: int) -> int:
    return load().oracle_pcg(v)


def owen(x: int, seed: int) -> int:
    return load().oracle_owen(x, seed)


def sobol2(i: int) -> int:
    return load().oracle_sobol2(i)


def reverse_bits(x: int) -> int:
    return load().oracle_reverse_bits(x)


def qmc_draws(pixel: int, sample: int, dim0: int, n: int, pairs: bool) -> np.ndarray:
    out = np.zeros(n * (2 if pairs else 1), np.float32)
    load().oracle_qmc_draws(pixel, sample, dim0, n, int(pairs), _p(out))
    return out.reshape(n, 2) if pairs else out


def eval_bsdf(material_words: np.ndarray, base, n, wo, wi) -> np.ndarray:
    out = np.zeros(4, np.float32)
    args = [np.ascontiguousarray(x, np.float32) for x in (base, n, wo, wi)]
    load().oracle_eval_bsdf(_p(np.ascontiguousarray(material_words, np.uint32)), *[_p(a) for a in args], _p(out))
    return out


def ggx_terms(roughness: float, nh: float, nv: float, nl: float) -> np.ndarray:
    """(D(n·h), G1(n·v), G1(n·l)) for a roughness."""
    out = np.zeros(3, np.float32)
    load().oracle_ggx_terms(roughness, nh, nv, nl, _p(out))
    return out


def ggx_sample(roughness: float, u1: float, u2: float) -> np.ndarray:
    out = np.zeros(3, np.float32)
    load().oracle_ggx_sample(roughness, u1, u2, _p(out))
    return out


def sample_bsdf(material_words: np.ndarray, base, n, wo, seed: int) -> np.ndarray:
    """(wi[3], weight[3], pdf, delta, ok) of one sample_bsdf draw."""
    out = np.zeros(9, np.float32)
    args = [np.ascontiguousarray(x, np.float32) for x in (base, n, wo)]
    load().oracle_sample_bsdf(_p(np.ascontiguousarray(material_words, np.uint32)), *[_p(a) for a in args], seed, _p(out))
    return out


def fresnel_dielectric(cos_i: float, eta: float) -> float:
    return load().oracle_fresnel_dielectric(cos_i, eta)


def wavelength_rgb(l: float) -> np.ndarray:
    out = np.zeros(3, np.float32)
    load().oracle_wavelength_rgb(l, _p(out))
    return out
