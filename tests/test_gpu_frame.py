"""The path-tracing kernel against the oracle under identical RNG streams.

With one sample per pixel, each pixel's accumulation is one radiance()
value (src/shaders.mjs:1179-1359), so the GPU and the oracle compute the same
path sample by sample. Paths only part where a transcendental (cos, exp, pow,
atan2 ...) differs by an ulp between the GPU's and the CPU's maths library and
flips a branch, so almost every pixel must agree to float rounding. Each case
exercises a feature: materials and NEE+MIS (spheres), glass and boxes
(cornell), volumes + environment map (triangles with fog), dispersion
(caustics), textures and normal maps, patterned lamps, environment NEE."""
import numpy as np
import pytest

from helpers import agreement, gpu_render, oracle_render

pytestmark = pytest.mark.gpu

CASES = [
    ("spheres", {"scene": {"width": 160, "height": 90}, "width": 160, "height": 90}),
    ("cornell", {"width": 96, "height": 96}),
    ("mixed", {"width": 96, "height": 96}),
    ("triangles", {"scene": {"grid": 48, "extras": 3, "envWidth": 128, "envHeight": 64}, "width": 96, "height": 54,
                   "maxBounces": 12}),
    ("caustics", {"scene": {"res": 24}, "width": 96, "height": 54}),
    ("fixture", {"scene": {"name": "textured", "size": 48}}),
    ("fixture", {"scene": {"name": "bumpy", "size": 48}}),
    ("fixture", {"scene": {"name": "patternedLamps", "size": 48}}),
    ("fixture", {"scene": {"name": "environmentFloor", "size": 48}}),
    ("fixture", {"scene": {"name": "furnaceDispersive", "size": 48}}),
    ("fixture", {"scene": {"name": "lampInFog", "size": 48}}),
    ("fixture", {"scene": {"name": "blackSmoke", "size": 48}}),
]


def case_id(c):
    return c[1]["scene"]["name"] if c[0] == "fixture" else c[0]


@pytest.mark.parametrize("config,opts", CASES, ids=[case_id(c) for c in CASES])
def test_first_frame_per_sample(ctx, pack, config, opts):
    p = pack(config, **opts)
    params = p.with_params(caustics=0, denoise=0, firefly_clamp=0.0)
    gpu = gpu_render(ctx, p, 1, params)
    cpu = oracle_render(p, 1, params)
    for name, g, c in zip(("accum", "albedo", "normal"), gpu, cpu):
        assert np.isfinite(g).all(), name
        frac = agreement(g, c)
        print(f"\n{case_id((config, opts))} {name}: {frac:.5f} of pixels agree")
        assert frac >= 0.999, f"{case_id((config, opts))} {name}: only {frac:.4f} of pixels agree"


@pytest.mark.parametrize("config,opts", [CASES[0], CASES[3], CASES[5]], ids=["spheres", "triangles", "textured"])
def test_later_frames_accumulate(ctx, pack, config, opts):
    """Four frames: sample offsets 0..3 draw different Sobol points, and frames after the first add."""
    p = pack(config, **opts)
    params = p.with_params(caustics=0, denoise=0, firefly_clamp=0.0)
    gpu = gpu_render(ctx, p, 4, params)
    cpu = oracle_render(p, 4, params)
    frac = agreement(gpu[0], cpu[0], rel=1e-4, abs_=1e-4)
    print(f"\n{config} after 4 frames: {frac:.5f} of pixels agree")
    assert frac >= 0.999, f"{config}: only {frac:.4f} of pixels agree after 4 frames"


@pytest.mark.parametrize("view", [1, 2, 3], ids=["albedo", "normal", "depth"])
@pytest.mark.parametrize("config,opts", [CASES[2], CASES[6]], ids=["mixed", "bumpy"])
def test_inspection_views(ctx, pack, config, opts, view):
    """Inspection views trace the pixel centre once (src/shaders.mjs:1377-1407): no randomness."""
    p = pack(config, **opts)
    params = p.with_params(caustics=0, denoise=0, view=view)
    g = gpu_render(ctx, p, 1, params)[0]
    c = oracle_render(p, 1, params)[0]
    frac = agreement(g, c, rel=2e-6, abs_=1e-6)
    assert frac > 0.995, f"view {view}: {frac:.4f}"
    if view == 3:  # depth: the hit distance itself, bit-exact
        assert np.array_equal(g[..., 0].view(np.uint32), c[..., 0].view(np.uint32))


def test_firefly_clamp(ctx, pack):
    p = pack(*CASES[0][:1], **CASES[0][1])
    params = p.with_params(caustics=0, denoise=0, firefly_clamp=1.5)
    g = gpu_render(ctx, p, 1, params)[0]
    assert g[..., :3].max() <= 1.5 * (1 + 1e-6)
    c = oracle_render(p, 1, params)[0]
    assert agreement(g, c) >= 0.999


@pytest.mark.parametrize("config,opts,spp", [(CASES[0][0], CASES[0][1], 1), (CASES[0][0], CASES[0][1], 3),
                                              (CASES[3][0], CASES[3][1], 2), (CASES[4][0], CASES[4][1], 1),
                                              (CASES[5][0], CASES[5][1], 2)],
                         ids=["spheres", "spheres-3spp", "triangles-fog-env", "dispersive", "textured"])
def test_schedules_are_bit_identical(ctx, pack, config, opts, spp):
    """The compacting, lane-regenerating, wavefront and split schedules, and
    the one-thread-per-pixel pass compiled for five waves per SIMD, compute
    the same samples, each pixel's in the same order, as one thread per pixel
    at four: accumulation and guides equal bit for bit. The split schedule
    runs at three split depths and at five waves (at 3 spp a frame, k_main
    renders for it)."""
    from raymond_amd import capi
    p = pack(config, **opts)
    params = p.with_params(caustics=0, denoise=0, firefly_clamp=0.0, samples_per_frame=spp)
    try:
        ctx.set_option(capi.OPTION_SCHEDULE, capi.SCHEDULE_PIXEL)
        plain = gpu_render(ctx, p, 3, params)
        runs = [(capi.SCHEDULE_PIXEL, 0, 5, 8, 0, 0), (capi.SCHEDULE_PIXEL, 0, 4, 16, 0, 0),
                (capi.SCHEDULE_PIXEL, 0, 4, 8, 1, 0), (capi.SCHEDULE_COMPACT, 0, 4, 8, 0, 0),
                (capi.SCHEDULE_SORTED, 0, 4, 8, 0, 0), (capi.SCHEDULE_PERSISTENT, 0, 4, 8, 0, 0),
                (capi.SCHEDULE_PERSISTENT, 7, 4, 8, 0, 0), (capi.SCHEDULE_WAVEFRONT, 0, 4, 8, 0, 0),
                (capi.SCHEDULE_SPLIT, 0, 4, 8, 0, 1), (capi.SCHEDULE_SPLIT, 0, 4, 8, 0, 2),
                (capi.SCHEDULE_SPLIT, 0, 4, 8, 0, 3), (capi.SCHEDULE_SPLIT, 0, 5, 8, 0, 2)]
        for schedule, blocks, waves, depth, shape, split in runs:
            ctx.set_option(capi.OPTION_SCHEDULE, schedule)
            ctx.set_option(capi.OPTION_PERSISTENT_BLOCKS, blocks)
            ctx.set_option(capi.OPTION_WAVES, waves)
            ctx.set_option(capi.OPTION_STACK_DEPTH, depth)
            ctx.set_option(capi.OPTION_BLOCK_SHAPE, shape)
            ctx.set_option(capi.OPTION_SPLIT_DEPTH, split)
            other = gpu_render(ctx, p, 3, params)
            for name, a, b in zip(["accum", "albedo", "normal"], plain, other):
                differ = int(np.any(a.view(np.uint32) != b.view(np.uint32), axis=-1).sum())
                assert differ == 0, (f"schedule {schedule} (blocks {blocks}, waves {waves}, stack {depth}, "
                                     f"shape {shape}, split depth {split}): {differ} pixels of {name} differ")
    finally:
        ctx.set_option(capi.OPTION_SPLIT_DEPTH, 0)
        ctx.set_option(capi.OPTION_SCHEDULE, capi.SCHEDULE_DEFAULT)
        ctx.set_option(capi.OPTION_PERSISTENT_BLOCKS, 0)
        ctx.set_option(capi.OPTION_WAVES, 0)
        ctx.set_option(capi.OPTION_STACK_DEPTH, 0)
        ctx.set_option(capi.OPTION_BLOCK_SHAPE, 0)
