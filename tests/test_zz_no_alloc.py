"""rm_raycast_static, the raycast entry point that allocates nothing (its
traversal spill is a static device array, its rays and hits device memory
the caller owns), against rm_raycast on the same rays: the same hit records
and traversal counters bit for bit, on two calls in a row, the second with
fewer rays, through the FP32 and the quantized tree."""
import ctypes as C

import numpy as np
import pytest

from raymond_amd import capi
from test_gpu_trace import random_rays

pytestmark = pytest.mark.gpu

# Deep enough that traversal stacks pass the 8 entries kept in LDS, so the
# static spill is used: a 256² fBm terrain with tori, lathes and tubes.
SCENE = ("triangles", {"scene": {"grid": 256, "extras": 6, "envWidth": 16, "envHeight": 8}, "width": 64,
                       "height": 36})
H2D, D2H = 1, 2


class DeviceMemory:
    """Device buffers from the HIP runtime the library itself uses (not the library)."""

    def __init__(self):
        capi.load()
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        self.held = []

    def alloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), nbytes) == 0
        self.held.append(p.value)
        return p.value

    def put(self, a: np.ndarray) -> int:
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def get(self, p: int, shape, dtype) -> np.ndarray:
        out = np.empty(shape, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def poison(self, p: int, nbytes: int) -> None:
        assert self.hip.hipMemset(p, 0xAB, nbytes) == 0

    def free(self):
        for p in self.held:
            self.hip.hipFree(p)
        self.held = []


@pytest.mark.parametrize("precision", [0, capi.PRECISION_QUANTIZED_BVH], ids=["fp32", "quantized"])
def test_static_raycast_equals_raycast(ctx, pack, precision):
    p = pack(*SCENE[:1], **SCENE[1])
    big = random_rays(65536, seed=41, spread=25.0, target=12.0)
    small = random_rays(1000, seed=43, spread=25.0, target=12.0)
    mem = DeviceMemory()
    scene = ctx.upload_scene(p)
    ctx.set_option(capi.OPTION_PRECISION, precision)
    try:
        rays_dev = mem.put(big)
        hits_dev = mem.alloc(65536 * 16)
        cnt_dev = mem.alloc(65536 * 8)
        for rays, any_hit in ((big, False), (small, True)):
            n = len(rays)
            if rays is small:
                assert mem.hip.hipMemcpy(rays_dev, small.ctypes.data, small.nbytes, H2D) == 0
            mem.poison(hits_dev, 65536 * 16)
            mem.poison(cnt_dev, 65536 * 8)
            ctx.raycast_static(scene, rays_dev, n, hits_dev, any_hit=any_hit, counters=cnt_dev)
            ctx.synchronize()
            got = mem.get(hits_dev, (n, 4), np.uint32)
            got_cnt = mem.get(cnt_dev, (n, 2), np.uint32)
            want, want_cnt = ctx.raycast(scene, rays, any_hit=any_hit, counters=True)
            assert (want[:, 1] != 0xFFFFFFFF).mean() > 0.3
            np.testing.assert_array_equal(got, want)
            np.testing.assert_array_equal(got_cnt, want_cnt)
            # Beyond n nothing is written.
            rest = mem.get(hits_dev, (65536, 4), np.uint32)[n:]
            assert (rest == 0xABABABAB).all()
        print(f"\nstatic raycast: deepest traversal tested {int(want_cnt[:, 0].max())} node records")
        with pytest.raises(capi.RaymondError, match="at most 65536 rays"):
            ctx.raycast_static(scene, rays_dev, 65537, hits_dev)
    finally:
        ctx.set_option(capi.OPTION_PRECISION, 0)
        scene.release()
        mem.free()
