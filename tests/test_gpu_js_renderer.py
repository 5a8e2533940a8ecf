"""The JS Renderer end to end on the GPU: Renderer.create → render() ×4 →
readPixels (average and denoised) → readDisplay → dispose, through
raymond.node. Its average must equal, bit for bit, the same frames driven
through the C ABI from the same packed buffers, since both run the same
kernels on the same inputs; the PNG of its display pass decodes to the
oracle's display of those buffers."""
import json
import os
import struct
import subprocess
import zlib
from pathlib import Path

import numpy as np
import pytest

from helpers import frame_params, gpu_render
from oracle import pyoracle
from raymond_amd import capi

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent


def test_renderer_matches_c_abi(ctx, pack, tmp_path):
    frames = 4
    res = subprocess.run(["node", str(ROOT / "tests" / "js" / "renderer_gpu.mjs"), str(tmp_path / "avg.bin"),
                          str(tmp_path / "den.bin"), str(frames), str(tmp_path / "shown.png")],
                         capture_output=True, text=True, timeout=120,
                         env=dict(os.environ, RAYMOND_EXPECT_GPU="1"))
    assert res.returncode == 0, res.stdout + res.stderr
    assert json.loads(res.stdout.strip().splitlines()[-1])["samples"] == frames
    avg = np.fromfile(tmp_path / "avg.bin", np.float32).reshape(64, 64, 4)
    den = np.fromfile(tmp_path / "den.bin", np.float32).reshape(64, 64, 4)
    p = pack("cornell", width=64, height=64, caustics=False)
    params = p.with_params(firefly_clamp=0.0, caustics=0, denoise=1)
    acc, _, _, d = gpu_render(ctx, p, frames, params, denoised=True)
    np.testing.assert_array_equal(avg, acc / frames)
    np.testing.assert_array_equal(den, d)
    shown = read_png(tmp_path / "shown.png")
    expect = pyoracle.display(frame_params(p, frames - 1, params), acc, d)
    diff = np.abs(shown.astype(int) - expect.astype(int))
    assert diff.max() <= 1 and (diff == 0).mean() > 0.99, diff.max()


def read_png(path) -> np.ndarray:
    """An 8-bit RGBA PNG of unfiltered rows, as encodePng writes it, as an (h, w, 4) array."""
    data = Path(path).read_bytes()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    at, idat = 8, b""
    while at < len(data):
        (n,) = struct.unpack(">I", data[at:at + 4])
        kind, body = data[at + 4:at + 8], data[at + 8:at + 8 + n]
        assert struct.unpack(">I", data[at + 8 + n:at + 12 + n])[0] == zlib.crc32(kind + body)
        if kind == b"IHDR":
            w, h, depth, color = struct.unpack(">IIBB", body[:10])
            assert (depth, color) == (8, 6)
        elif kind == b"IDAT":
            idat += body
        at += 12 + n
    rows = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, w * 4 + 1)
    assert (rows[:, 0] == 0).all()
    return rows[:, 1:].reshape(h, w, 4)


def test_render_server_matches_renderer():
    """RenderServer jobs (one in a mix, one a copy, one without the denoiser)
    give the Renderer's images for the same frames, bit for bit."""
    res = subprocess.run(["node", str(ROOT / "tests" / "js" / "server_gpu.mjs")], capture_output=True, text=True,
                         timeout=180, env=dict(os.environ, RAYMOND_EXPECT_GPU="1"))
    assert res.returncode == 0, res.stdout + res.stderr
    out = json.loads(res.stdout.strip().splitlines()[-1])
    assert out["denoisedFirst"] and out["denoisedSecond"] and out["average"], out
    for st in out["stats"]:
        assert 0 <= st["queuedMs"] <= st["firstFrameMs"] <= st["doneMs"], st
    assert [st["samples"] for st in out["stats"]] == [5, 3, 5, 5]
