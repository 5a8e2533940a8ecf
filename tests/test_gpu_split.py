"""The split schedule, which contexts start with: a thread a pixel for the
first RM_OPTION_SPLIT_DEPTH bounces, then persistent waves over the queue of
the paths still alive. Against one thread per pixel, bit for bit, where the
queue and the tail's grid take their edge shapes: images smaller than a wave
or a block and sizes that leave partial 16×16 tiles (the tail's grid is cut
to what the queue could fill); bounce limits at or below the split depth
(nothing queued) and the deepest split; the quantized tree, compensated sums
and the inspection views (which the per-pixel pass renders); and the
option's argument check."""
import numpy as np
import pytest

from helpers import gpu_render
from raymond_amd import capi

pytestmark = pytest.mark.gpu


def render_with(ctx, p, params, schedule, depth=0, frames=2, precision=0):
    ctx.set_option(capi.OPTION_SCHEDULE, schedule)
    ctx.set_option(capi.OPTION_SPLIT_DEPTH, depth)
    ctx.set_option(capi.OPTION_PRECISION, precision)
    try:
        return gpu_render(ctx, p, frames, params)
    finally:
        ctx.set_option(capi.OPTION_SCHEDULE, capi.SCHEDULE_DEFAULT)
        ctx.set_option(capi.OPTION_SPLIT_DEPTH, 0)
        ctx.set_option(capi.OPTION_PRECISION, 0)


def assert_same(got, want, what):
    for name, a, b in zip(("accum", "albedo", "normal"), got, want):
        differ = int(np.any(a.view(np.uint32) != b.view(np.uint32), axis=-1).sum())
        assert differ == 0, f"{what}: {differ} pixels of {name} differ"


@pytest.mark.parametrize("w,h", [(1, 1), (7, 3), (17, 5), (257, 3), (64, 64)])
def test_split_matches_per_pixel_at_edge_sizes(ctx, pack, w, h):
    p = pack("spheres", scene={"width": w, "height": h}, width=w, height=h)
    params = p.with_params(caustics=0, denoise=0, firefly_clamp=0.0)
    assert_same(render_with(ctx, p, params, capi.SCHEDULE_SPLIT),
                render_with(ctx, p, params, capi.SCHEDULE_PIXEL), f"{w}x{h}")


@pytest.mark.parametrize("bounces,depth", [(0, 2), (1, 2), (2, 2), (8, 8), (8, 1), (3, 3)])
def test_split_matches_per_pixel_at_bounce_limits(ctx, pack, bounces, depth):
    p = pack("cornell", width=64, height=64)
    params = p.with_params(caustics=0, denoise=0, firefly_clamp=0.0, max_bounces=bounces)
    assert_same(render_with(ctx, p, params, capi.SCHEDULE_SPLIT, depth),
                render_with(ctx, p, params, capi.SCHEDULE_PIXEL), f"{bounces} bounces, split after {depth}")


@pytest.mark.parametrize("precision", [capi.PRECISION_QUANTIZED_BVH, capi.PRECISION_COMPENSATED],
                         ids=["quantized-bvh", "compensated"])
def test_split_matches_per_pixel_with_precision_variants(ctx, pack, precision):
    p = pack("mixed", width=80, height=48)
    params = p.with_params(caustics=0, denoise=0, firefly_clamp=0.0)
    assert_same(render_with(ctx, p, params, capi.SCHEDULE_SPLIT, precision=precision, frames=3),
                render_with(ctx, p, params, capi.SCHEDULE_PIXEL, precision=precision, frames=3),
                f"precision {precision}")


@pytest.mark.parametrize("view", [1, 2, 3], ids=["albedo", "normal", "depth"])
def test_split_renders_the_views_as_per_pixel(ctx, pack, view):
    p = pack("cornell", width=48, height=48)
    params = p.with_params(caustics=0, denoise=0, firefly_clamp=0.0, view=view)
    assert_same(render_with(ctx, p, params, capi.SCHEDULE_SPLIT, frames=1),
                render_with(ctx, p, params, capi.SCHEDULE_PIXEL, frames=1), f"view {view}")


def test_split_depth_is_checked(ctx):
    for bad in (-1, 9):
        with pytest.raises(capi.RaymondError, match="SPLIT_DEPTH"):
            ctx.set_option(capi.OPTION_SPLIT_DEPTH, bad)
