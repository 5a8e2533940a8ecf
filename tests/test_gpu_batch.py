"""Cross-job batching (rm_frame_batch, and the job server under
RM_OPTION_BATCH): several contexts' path passes in one launch of the
per-pixel kernel, or, with every context on the wavefront schedule, each
wavefront stage one launch for all the jobs, the launch sequence replayed
from a hipGraph; the light tracing of the frames that have it in one launch;
the denoiser's à-trous passes and compose of the frames that denoise in one
launch each. Each job's frame is computed by the same code on the same
inputs as alone, so its sums and denoised image are the same bit for bit,
even where the batch's kernel carries features (volumes, environment,
textures) the job lacks; the light-traced jobs agree to the light tests'
tolerance (photons land in a cell in the order their atomics do). The
wavefront batch is checked after every frame, so every replay of its graph
is compared with frames rendered alone, with the program's own launches and
read-backs between the replays."""
import numpy as np
import pytest

from helpers import agreement, frame_params, gpu_render
from raymond_amd import capi, scenefile

pytestmark = pytest.mark.gpu

JOBS = [
    ("spheres", {"scene": {"width": 96, "height": 54}, "width": 96, "height": 54}, dict(caustics=0, denoise=1)),
    ("triangles", {"scene": {"grid": 32, "extras": 3, "envWidth": 64, "envHeight": 32, "width": 80, "height": 48},
                   "width": 80, "height": 48}, dict(caustics=0, denoise=0)),
    ("cornell", {"width": 64, "height": 64}, dict(caustics=0, denoise=1, samples_per_frame=2)),
    ("fixture", {"scene": {"name": "textured", "size": 48}}, dict(caustics=0, denoise=0)),
    ("fixture", {"scene": {"name": "causticLamp", "size": 48}}, dict(caustics=1, denoise=0)),
    ("fixture", {"scene": {"name": "causticLamp", "size": 40}}, dict(caustics=1, denoise=1)),
]
FRAMES = 3


def check_against_alone(got, alone, when):
    for (name, _, f), a, g in zip(JOBS, alone, got):
        if f.get("caustics"):
            assert agreement(g[0], a[0], rel=1e-3, abs_=1e-4) > 0.97, f"{name} {when}"
            continue
        for what, x, y in zip(("accum", "albedo", "normal", "denoised"), g, a):
            if what == "denoised" and not f.get("denoise"):
                continue
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), f"{name}: {what} {when}"


@pytest.mark.parametrize("schedule", ["per-pixel", "wavefront"])
def test_batched_frames_match_frames_rendered_alone(ctx, pack, schedule):
    packs = [pack(c, **o) for c, o, _ in JOBS]
    params = [p.with_params(firefly_clamp=0.0, **f) for p, (_, _, f) in zip(packs, JOBS)]
    # Alone: the per-pixel pass (the wavefront pass gives the same sums).
    alone = [[gpu_render(ctx, p, k + 1, q, denoised=True) for k in range(FRAMES)] for p, q in zip(packs, params)]
    contexts = [capi.Context(0) for _ in JOBS]
    scenes = []
    try:
        for c, p in zip(contexts, packs):
            scenes.append(c.upload_scene(p))
            c.upload_environment(p)
            c.resize(p.width, p.height)
            if schedule == "wavefront":
                c.set_option(capi.OPTION_SCHEDULE, capi.SCHEDULE_WAVEFRONT)
        for k in range(FRAMES):
            fp = [frame_params(p, k, q) for p, q in zip(packs, params)]
            em = [p.emitters if scenefile.get_param(q, "caustics") else None for p, q in zip(packs, params)]
            capi.frame_batch(contexts, scenes, fp, em)
            if schedule == "wavefront" or k == FRAMES - 1:
                got = [[c.read(b) for b in (capi.BUFFER_ACCUM, capi.BUFFER_GUIDE_ALBEDO, capi.BUFFER_GUIDE_NORMAL,
                                            capi.BUFFER_DENOISED)] for c in contexts]
                check_against_alone(got, [a[k] for a in alone], f"after frame {k + 1}")
    finally:
        for c, s in zip(contexts, scenes):
            s.release()
        for c in contexts:
            c.close()


def test_batch_arguments_are_checked(ctx, pack):
    p = pack(*JOBS[0][:1], **JOBS[0][1])
    q = p.with_params(caustics=0, denoise=0)
    scene = ctx.upload_scene(p)
    other = capi.Context(0)
    try:
        ctx.resize(p.width, p.height)
        with pytest.raises(capi.RaymondError, match="one frame per context"):
            capi.frame_batch([ctx, ctx], [scene, scene], [q, q])
        with pytest.raises(capi.RaymondError, match="1 to 8"):
            capi.frame_batch([ctx] * 9, [scene] * 9, [q] * 9)
        other.resize(p.width, p.height)
        other.set_option(capi.OPTION_PRECISION, capi.PRECISION_QUANTIZED_BVH)
        with pytest.raises(capi.RaymondError, match="same kind of tree"):
            capi.frame_batch([ctx, other], [scene, scene], [q, q])
    finally:
        scene.release()
        other.close()


@pytest.mark.parametrize("schedule", ["per-pixel", "wavefront"])
def test_server_batches_jobs_that_match_their_frames_alone(ctx, pack, schedule):
    chosen = [j for j in JOBS if not j[2].get("caustics")]
    packs = [pack(c, **o) for c, o, _ in chosen]
    params = [p.with_params(firefly_clamp=0.0, **f) for p, (_, _, f) in zip(packs, chosen)]
    expected = []
    for p, q, (_, _, f) in zip(packs, params, chosen):
        out = gpu_render(ctx, p, FRAMES, q, denoised=True)
        expected.append(out[3] if f.get("denoise") else out[0])
    with capi.Server(slots=4, memory_budget=8 << 30) as server:
        server.set_option(capi.OPTION_BATCH, 1)
        if schedule == "wavefront":
            server.set_option(capi.OPTION_SCHEDULE, capi.SCHEDULE_WAVEFRONT)
        scenes = [server.context.upload_scene(p) for p in packs]
        spp = [int(q.view(np.uint32)[43]) for q in params]
        ids = [server.submit(s, q, FRAMES * k, packed=p) for p, q, k, s in zip(packs, params, spp, scenes)]
        ids += [server.submit(s, q, FRAMES * k, packed=p) for p, q, k, s in zip(packs, params, spp, scenes)]
        got = [server.wait(i) for i in ids]
    for k, (image, stats) in enumerate(got):
        j = k % len(chosen)
        assert stats["frames"] == FRAMES, chosen[j][0]
        assert np.array_equal(image.view(np.uint32), expected[j].view(np.uint32)), chosen[j][0]
