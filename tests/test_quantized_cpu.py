"""The 8-bit child-quantized tree (rm_bvh_quantize) on the host: decoded in
FP32 exactly as the device decodes it, every child box encloses its FP32
box, so no ray that reaches a leaf in the FP32 tree misses it in the
quantized one; leaf records carry the leaf's range; trees that can't be
enclosed are refused."""
import numpy as np
import pytest

from raymond_amd import capi

SCENES = [
    ("cornell", {"width": 32, "height": 32}),
    ("spheres", {"scene": {"width": 32, "height": 32}, "width": 32, "height": 32}),
    ("mixed", {"width": 32, "height": 32}),
    ("triangles", {"scene": {"grid": 48, "extras": 3, "envWidth": 16, "envHeight": 8}, "width": 32, "height": 32}),
    ("caustics", {"scene": {"res": 24}, "width": 32, "height": 32}),
]


def decode(q):
    """Child boxes (2 per record: lo, hi as (N, 2, 2, 3) float32) and meta of quantized records."""
    f = q.view(np.float32)
    origin = f[:, :3]
    meta = q[:, 3]
    step = np.stack([((meta >> (8 * a)) & 0xFF).astype(np.uint32) << 23 for a in range(3)], 1).view(np.float32)
    b = q[:, 4:7].copy().view(np.uint8).reshape(-1, 12).astype(np.float32)   # little-endian bytes in order
    codes = b.reshape(-1, 2, 2, 3)   # child, lo/hi, axis
    boxes = origin[:, None, None, :] + codes * step[:, None, None, :]   # float32: q·step exact, one rounding
    return boxes, meta


@pytest.mark.parametrize("config,opts", SCENES, ids=[s[0] for s in SCENES])
def test_decoded_children_enclose_the_fp32_boxes(pack, config, opts):
    nodes = pack(config, **opts).nodes.reshape(-1, 8)
    q = capi.bvh_quantize(nodes)
    f = nodes.view(np.float32)
    inner = np.nonzero(nodes[:, 7] == 0)[0]
    leaves = np.nonzero(nodes[:, 7] > 0)[0]
    np.testing.assert_array_equal(q[leaves, :2], nodes[leaves][:, [3, 7]])
    assert not q[leaves, 2:].any()
    boxes, meta = decode(q[inner])
    first = nodes[inner, 3]
    np.testing.assert_array_equal(q[inner, 7], first)
    np.testing.assert_array_equal(q[inner].view(np.float32)[:, :3], f[inner, :3])
    assert inner.size > 0
    inflation = []
    for k in (0, 1):
        child = first + k
        lo, hi = f[child, :3], f[child, 4:7]
        assert (boxes[:, k, 0] <= lo).all() and (boxes[:, k, 1] >= hi).all(), "a decoded box cuts its child"
        np.testing.assert_array_equal((meta >> (24 + k)) & 1, (nodes[child, 7] > 0).astype(np.uint32))
        extent = np.maximum(f[inner, 4:7] - f[inner, :3], 1e-30)
        inflation.append(((lo - boxes[:, k, 0]) + (boxes[:, k, 1] - hi)) / extent)
    # Outward rounding costs at most a step (parent extent / 255, doubled at worst by the exponent) per side.
    assert np.max(inflation) <= 4.0 * 2 / 255 + 1e-6


def test_refuses_what_it_cannot_enclose():
    nodes = np.zeros((3, 8), np.uint32)
    f = nodes.view(np.float32)
    f[0, :3], f[0, 4:7] = 0.0, 1.0
    nodes[0, 3], nodes[0, 7] = 1, 0
    for k in (1, 2):
        f[k, :3], f[k, 4:7] = 0.25, 0.5
        nodes[k, 3], nodes[k, 7] = k - 1, 1
    q = capi.bvh_quantize(nodes)
    boxes, _ = decode(q[:1])
    assert (boxes[0, :, 0] <= 0.25).all() and (boxes[0, :, 1] >= 0.5).all()
    f[2, 4] = 2.0                      # a child outside its parent
    with pytest.raises(capi.RaymondError, match="outside its parent"):
        capi.bvh_quantize(nodes)
    f[2, 4] = 0.5
    f[0, 4] = np.inf                   # a box that isn't finite
    with pytest.raises(capi.RaymondError, match="finite"):
        capi.bvh_quantize(nodes)
    nodes[0, 3] = 2                    # children past the end
    f[0, 4] = 1.0
    with pytest.raises(capi.RaymondError, match="outside the tree"):
        capi.bvh_quantize(nodes)


def decode_wide(w):
    """Wide records: child boxes (W, 4, 2, 3) float32, meta, refs, counts."""
    f = w.view(np.float32)
    origin = f[:, :3]
    meta = w[:, 3]
    step = np.stack([((meta >> (8 * a)) & 0xFF).astype(np.uint32) << 23 for a in range(3)], 1).view(np.float32)
    codes = w[:, 4:10].copy().view(np.uint8).reshape(-1, 4, 2, 3).astype(np.float32)
    boxes = origin[:, None, None, :] + codes * step[:, None, None, :]
    counts = w[:, 14].copy().view(np.uint8).reshape(-1, 4)
    return boxes, meta, w[:, 10:14], counts


@pytest.mark.parametrize("config,opts", SCENES, ids=[s[0] for s in SCENES])
def test_wide_tree_covers_every_leaf_and_encloses_its_boxes(pack, config, opts):
    """Walking the wide records from the root reaches every leaf of the binary
    tree exactly once, each decoded box enclosing the FP32 box of the binary
    node it stands for."""
    nodes = pack(config, **opts).nodes.reshape(-1, 8)
    f = nodes.view(np.float32)
    w = capi.bvh_quantize_wide(nodes)
    if nodes[0, 7] > 0:
        assert w.shape[0] == 0
        return
    boxes, meta, refs, counts = decode_wide(w)
    # The binary nodes each wide child stands for, by the fold's rule.
    def kids(p):
        out = []
        for k in (nodes[p, 3], nodes[p, 3] + 1):
            out += [k] if nodes[k, 7] > 0 else [nodes[k, 3], nodes[k, 3] + 1]
        return out
    seen_leaves, todo, visited = [], [(0, 0)], 0
    while todo:
        wi, p = todo.pop()
        visited += 1
        ks = kids(p)
        assert meta[wi] >> 28 == len(ks)
        for slot, k in enumerate(ks):
            lo, hi = f[k, :3], f[k, 4:7]
            assert (boxes[wi, slot, 0] <= lo).all() and (boxes[wi, slot, 1] >= hi).all()
            leaf = bool(meta[wi] >> (24 + slot) & 1)
            assert leaf == (nodes[k, 7] > 0)
            if leaf:
                assert refs[wi, slot] == nodes[k, 3] and counts[wi, slot] == nodes[k, 7]
                seen_leaves.append(k)
            else:
                todo.append((refs[wi, slot], k))
    assert visited == w.shape[0]
    leaves = np.nonzero(nodes[:, 7] > 0)[0]
    assert sorted(seen_leaves) == sorted(leaves.tolist())
