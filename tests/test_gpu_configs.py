"""The workload configurations at the sizes they are described at, on the GPU
against the oracle:

* the 1080p 200-sphere frame exactly as the bench renders it (light tracing,
  photon gather and the denoiser on);
* the ~2.2M-triangle scene (1024² fBm terrain, tori, lathes, Bézier tubes,
  fog box, 2048×1024 environment): hit records of random rays and of 1080p
  primary rays bit-exact with equal traversal counters, through the FP32
  tree and the quantized one, and the first 12-bounce frame per sample;
* the caustics scene at its full size (512² water sheet, dispersive ball,
  lamp and sun): light-traced 1080p frames;
* a Poisson job trace through the job server with the precision variants
  on and the 2.2M-triangle scene in its pool: every job's image equals the
  same job rendered alone.

The oracle runs the camera paths of one 54-row band of a 1080p frame (light
tracing and the gather over the whole image, since splats land anywhere), so
the GPU renders the whole frame and the band is compared."""
import numpy as np
import pytest

from helpers import agreement, frame_params, gpu_render, oracle_render
from oracle.pyoracle import Oracle
from raymond_amd import capi
from raymond_amd.scenefile import get_param, resized_params, set_params
from test_gpu_quantized import assert_same_closest
from test_gpu_trace import assert_same_hits, random_rays

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]

BAND = (513, 567)   # 54 rows through the middle of a 1080p frame
SPHERES = ("spheres", {})
TRIANGLES = ("triangles", {})
CAUSTICS = ("caustics", {})
# The triangle scene is 40 m across: rays from a 60 m cube aimed into the middle 30 m.
TRI_RAYS = dict(spread=30.0, target=15.0)


def band(a):
    return a[BAND[0]:BAND[1]]


def report(what, frac):
    print(f"\n{what}: {frac:.5f} of pixels agree")


# ------------------------------------------------------------------ C3: 2.2M triangles

def test_triangles_scene_is_the_described_size(pack):
    p = pack(*TRIANGLES[:1], **TRIANGLES[1])
    assert p.primitive_count > 2_100_000
    assert (p.width, p.height) == (1920, 1080)
    assert get_param(p.params, "max_bounces") == 12
    assert get_param(p.params, "volume_count") == 1
    assert (p.meta["envWidth"], p.meta["envHeight"]) == (2048, 1024)


@pytest.mark.parametrize("rays_of", ["random", "primary"])
def test_triangles_hits_bit_exact(ctx, pack, rays_of):
    """Closest and any-hit records bit for bit, and the same node and
    primitive records tested per ray, on 200k random rays or every 7th
    primary ray of the 1080p camera."""
    p = pack(*TRIANGLES[:1], **TRIANGLES[1])
    o = Oracle(p)
    rays = random_rays(200_000, seed=31, **TRI_RAYS) if rays_of == "random" else o.camera_rays()[::7]
    scene = ctx.upload_scene(p)
    try:
        for any_hit in (False, True):
            gpu, gcnt = ctx.raycast(scene, rays, any_hit=any_hit, counters=True)
            cpu, ccnt = o.trace(rays, any_hit=any_hit, counters=True)
            assert_same_hits(gpu, cpu, f"triangles {rays_of} rays any_hit={any_hit}")
            np.testing.assert_array_equal(gcnt, ccnt)
    finally:
        scene.release()
    hit = (cpu[:, 1] != 0xFFFFFFFF).mean()
    print(f"\ntriangles {rays_of}: {len(rays)} rays, {hit:.3f} hit; mean nodes {ccnt[:, 0].mean():.1f}, "
          f"primitives {ccnt[:, 1].mean():.1f} per any-hit ray")
    assert hit > 0.3


@pytest.mark.parametrize("rays_of", ["random", "primary"])
def test_triangles_quantized_hits(ctx, pack, rays_of):
    """Through the 4-wide quantized tree: the same hit records as the FP32
    oracle (nearest t, primitive, barycentrics) and the same occlusion."""
    p = pack(*TRIANGLES[:1], **TRIANGLES[1])
    o = Oracle(p)
    rays = random_rays(200_000, seed=37, **TRI_RAYS) if rays_of == "random" else o.camera_rays()[::7]
    scene = ctx.upload_scene(p)
    try:
        ctx.set_option(capi.OPTION_PRECISION, capi.PRECISION_QUANTIZED_BVH)
        gpu = ctx.raycast(scene, rays)
        occluded = ctx.raycast(scene, rays, any_hit=True)[:, 1] != 0xFFFFFFFF
    finally:
        ctx.set_option(capi.OPTION_PRECISION, 0)
        scene.release()
    cpu = o.trace(rays)
    ties = assert_same_closest(gpu, cpu, rays, o, f"triangles {rays_of}, quantized")
    assert ties == 0
    np.testing.assert_array_equal(occluded, o.trace(rays, any_hit=True)[:, 1] != 0xFFFFFFFF)


def test_triangles_first_frame_per_sample(ctx, pack):
    """The first 12-bounce frame (fog, environment, plastic) per sample on
    the middle band, under identical RNG streams."""
    p = pack(*TRIANGLES[:1], **TRIANGLES[1])
    params = p.with_params(caustics=0, denoise=0, firefly_clamp=0.0)
    gpu = gpu_render(ctx, p, 1, params)
    cpu = oracle_render(p, 1, params, rows=BAND)
    for name, g, c in zip(("accum", "albedo", "normal"), gpu, cpu):
        assert np.isfinite(g).all(), name
        frac = agreement(band(g), band(c))
        report(f"triangles 1080p band {name}", frac)
        assert frac >= 0.995, f"{name}: {frac:.5f}"


def test_triangles_quantized_frame_equals_fp32(ctx, pack):
    """The whole 1080p frame through the quantized tree equals the FP32
    tree's bit for bit: the same hits on every path."""
    p = pack(*TRIANGLES[:1], **TRIANGLES[1])
    params = p.with_params(caustics=0, denoise=0, firefly_clamp=0.0)
    plain = gpu_render(ctx, p, 1, params)
    ctx.set_option(capi.OPTION_PRECISION, capi.PRECISION_QUANTIZED_BVH)
    try:
        q = gpu_render(ctx, p, 1, params)
    finally:
        ctx.set_option(capi.OPTION_PRECISION, 0)
    for a, b, what in zip(q, plain, ("accum", "guide albedo", "guide normal")):
        differ = int(np.any(a.view(np.uint32) != b.view(np.uint32), axis=-1).sum())
        assert differ == 0, f"{what}: {differ} pixels differ"


# ------------------------------------------------------------------ C4: caustics at full size

def test_caustics_scene_is_the_described_size(pack):
    p = pack(*CAUSTICS[:1], **CAUSTICS[1])
    assert p.primitive_count == 2 * 511 * 511 + 3
    assert (p.width, p.height) == (1920, 1080)
    assert p.meta["causticsActive"]


def test_caustics_light_traced_frames(ctx, pack):
    """Two 1080p frames with light tracing (524,288 photons a frame, the
    reference's cap) and the photon gather: the middle band per pixel to
    the light tests' tolerance, and its total within 1%."""
    p = pack(*CAUSTICS[:1], **CAUSTICS[1])
    params = p.with_params(caustics=1, denoise=0, firefly_clamp=0.0)
    g = gpu_render(ctx, p, 2, params)[0]
    c = oracle_render(p, 2, params, rows=BAND)[0]
    assert np.isfinite(g).all()
    frac = agreement(band(g), band(c), rel=1e-3, abs_=1e-4)
    report("caustics 1080p band, 2 light-traced frames", frac)
    assert frac >= 0.9995
    tg, tc = band(g)[..., :3].sum(dtype=np.float64), band(c)[..., :3].sum(dtype=np.float64)
    assert abs(tg - tc) <= 0.01 * abs(tc), (tg, tc)


# ------------------------------------------------------------------ C2 as benched

def test_spheres_benched_frames(ctx, pack):
    """Four frames of the bench's workload as it renders them (1080p, 8
    bounces, light tracing, photon gather, denoiser): the sums of the middle
    band against the oracle, and the denoised band against the oracle's
    denoiser run on the GPU's own sums. The fourth frame's denoiser takes
    the fused pass (k_finish) with the temporal variance. The oracle's
    denoiser runs on a crop 80 rows beyond the band on each side, more than
    the 64-pixel reach of prepare and the five à-trous passes, so the band's
    denoised pixels do not see the crop's edge."""
    p = pack(*SPHERES[:1], **SPHERES[1])
    assert (p.width, p.height) == (1920, 1080) and p.primitive_count == 201
    assert p.meta["causticsActive"]
    frames = 4
    params = p.with_params(caustics=1, denoise=1, firefly_clamp=0.0)
    acc, ga, gn, den = gpu_render(ctx, p, frames, params, denoised=True)
    cpu = oracle_render(p, frames, params, rows=BAND)[0]
    assert np.isfinite(acc).all() and np.isfinite(den).all()
    frac = agreement(band(acc), band(cpu), rel=1e-3, abs_=1e-4)
    report("spheres200 1080p band, 4 benched frames (accum)", frac)
    assert frac >= 0.999

    margin = 80
    y0, y1 = BAND[0] - margin, BAND[1] + margin
    last = frame_params(p, frames - 1, params)
    crop = set_params(last.copy(), height=y1 - y0)
    passes = Oracle(p).denoise(crop, *(np.ascontiguousarray(a[y0:y1]) for a in (acc, ga, gn)))
    ref = passes[-1][margin:margin + BAND[1] - BAND[0]]
    frac = agreement(band(den), ref, rel=1e-4, abs_=1e-6)
    report("spheres200 1080p band, denoised", frac)
    assert frac > 0.995


# ------------------------------------------------------------------ C5: a served trace

def _alone(ctx, scene, p, params, frames):
    """A job rendered by itself on a context: frames of `params` at its own
    size, sample offsets, totals and photon radius advanced as the Renderer
    advances them; the denoised image when params.denoise is set."""
    w, h = get_param(params, "width"), get_param(params, "height")
    spp = get_param(params, "samples_per_frame")
    r0 = get_param(params, "photon_radius")
    caustics = get_param(params, "caustics")
    ctx.upload_environment(p)
    ctx.resize(w, h)
    for k in range(frames):
        r = r0 * (k + 1) ** (-1.0 / 6.0)
        fp = set_params(params.copy(), sample_offset=k * spp, total_samples=(k + 1) * spp, photon_radius=r,
                        photon_cell=2 * r)
        ctx.frame(scene, fp, p.emitters if caustics else None)
    return ctx.read(capi.BUFFER_DENOISED if get_param(params, "denoise") else capi.BUFFER_ACCUM)


def test_served_trace_matches_jobs_alone(ctx, pack):
    """Jobs of a Poisson trace (λ 20/s; scene pool from the 9-primitive
    Cornell box to the 2.2M-triangle scene; 256² to 1080p; 4 to 12 bounces)
    submitted at once to a server with 8 slots and a 200 GB budget, with
    the quantized tree, FP16 guides and compensated sums on: every image
    equals the job rendered alone with the same variants, bit for bit
    without light tracing, to the light tests' tolerance with it."""
    import bench
    trace = bench.serving_trace(rate=20.0, duration=0.75, spp_lo=2, spp_hi=6, seed=5)
    big = len(bench.SERVING_POOL) - 1
    # The largest scene at the largest size is in every trace of this test.
    trace.append((0.0, big, (1920, 1080), 3, 12))
    packs = [pack(c, **dict(o, maxBounces=12, samplesPerFrame=1, fireflyClamp=0.0, denoise=True, caustics=True))
             for c, o in bench.SERVING_POOL]
    assert packs[big].primitive_count > 2_100_000
    flags = capi.PRECISION_QUANTIZED_BVH | capi.PRECISION_FP16_GUIDES | capi.PRECISION_COMPENSATED
    jobs = []
    for _, k, (w, h), spp, bounces in trace:
        jobs.append((k, resized_params(packs[k].params, w, h, bounces), spp))

    with capi.Server(slots=8, memory_budget=200 << 30) as server:
        server.set_option(capi.OPTION_PRECISION, flags)
        scenes = [server.context.upload_scene(p) for p in packs]
        ids = []
        for k, params, spp in jobs:
            caustics = get_param(params, "caustics")
            ids.append(server.submit(scenes[k], params, spp, packs[k].emitters if caustics else None,
                                     packed=packs[k]))
        served = [server.wait(i) for i in ids]

    ctx.set_option(capi.OPTION_PRECISION, flags)
    uploaded = {}
    try:
        for (k, params, spp), (image, stats) in zip(jobs, served):
            assert stats["samples"] == spp and stats["frames"] == spp, stats
            if k not in uploaded:
                uploaded[k] = ctx.upload_scene(packs[k])
            expected = _alone(ctx, uploaded[k], packs[k], params, spp)
            what = f"job on pool scene {k} at {image.shape[1]}x{image.shape[0]}, {spp} spp"
            assert np.isfinite(image).all(), what
            if get_param(params, "caustics"):
                frac = agreement(image, expected, rel=1e-3, abs_=1e-4)
                report(what + " (light tracing)", frac)
                assert frac >= 0.999, what
            else:
                differ = int(np.any(image.view(np.uint32) != expected.view(np.uint32), axis=-1).sum())
                assert differ == 0, f"{what}: {differ} pixels differ"
    finally:
        ctx.set_option(capi.OPTION_PRECISION, 0)
        for s in uploaded.values():
            s.release()
    print(f"\nserved trace: {len(jobs)} jobs, pool scenes {sorted(set(k for k, _, _ in jobs))}")
