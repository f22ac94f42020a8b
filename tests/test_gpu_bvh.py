"""Invariants of the GPU LBVH (4-wide nodes): every triangle appears in
exactly one leaf, leaf boxes contain their triangles, child boxes nest in
their parent's entry, primitive ids form a permutation, depth is as
reported."""
import ctypes as C

import numpy as np
import pytest

from offlinerenderer_amd import scenes
from tests import gpu_util as G
from tests.scene_util import flatten

pytestmark = pytest.mark.gpu

_hip = None


def d2h(ptr, nbytes):
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    buf = np.empty(nbytes, dtype=np.uint8)
    _hip.hipDeviceSynchronize()
    assert _hip.hipMemcpy(buf.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), nbytes, 2) == 0
    return buf


def check_bvh(desc):
    ps = G.product_scene(desc)
    b = ps.getBVH()
    raw = d2h(b.nodes, b.numNodes * 128)
    f = raw.view(np.float32).reshape(b.numNodes, 32)
    u = raw.view(np.uint32).reshape(b.numNodes, 32)
    lo = np.stack([f[:, 0:4], f[:, 8:12], f[:, 16:20]], axis=2)   # node, child, axis
    hi = np.stack([f[:, 4:8], f[:, 12:16], f[:, 20:24]], axis=2)
    child = u[:, 24:28]
    tris = d2h(b.tris, b.numTris * 48).view(np.float32).reshape(b.numTris, 3, 4)
    prim = tris[:, 0, 3].view(np.uint32)
    V, F, _ = flatten(desc)
    assert np.array_equal(np.sort(prim), np.arange(b.numTris, dtype=np.uint32))
    assert np.allclose(tris[:, :, :3], V[F[prim]])
    covered = np.zeros(b.numTris, dtype=np.int32)
    seen = np.zeros(b.numNodes, dtype=np.int32)
    stack = [(0, 1)]
    maxd = 0
    while stack:
        node, depth = stack.pop()
        seen[node] += 1
        maxd = max(maxd, depth)
        for c in range(4):
            code = int(child[node, c])
            if code == 0xFFFFFFFF:
                continue
            blo, bhi = lo[node, c], hi[node, c]
            if code & 0x80000000:
                first, count = code & 0x07FFFFFF, ((code >> 27) & 0xF) + 1
                covered[first:first + count] += 1
                v = tris[first:first + count, :, :3].reshape(-1, 3)
                assert np.all(v >= blo) and np.all(v <= bhi)
            else:
                # the child's own children boxes nest inside this entry
                clo = lo[code][child[code] != 0xFFFFFFFF].min(0)
                chi = hi[code][child[code] != 0xFFFFFFFF].max(0)
                assert np.all(clo >= blo) and np.all(chi <= bhi)
                stack.append((code, depth + 1))
    assert np.all(covered == 1)
    assert np.all(seen == 1)
    assert maxd == b.maxDepth
    return b


def test_bvh_invariants_cornell(gpu):
    check_bvh(scenes.cornell(8, 8, 1))


def test_bvh_invariants_material(gpu):
    b = check_bvh(scenes.material_scene(32, 18, 1, tex_size=16))
    assert b.numNodes < b.numTris


def test_bvh_invariants_large(gpu):
    check_bvh(scenes.large_scene(8, 8, 1, num_spheres=60, num_emitters=4))


def test_bvh_tiny_scenes(gpu):
    for ntri in (1, 2, 3, 5):
        d = scenes.SceneDesc(width=4, height=4)
        d.ops.append(("lambert", (0.5, 0.5, 0.5), 0))
        rng = np.random.default_rng(ntri)
        v = rng.uniform(-1, 1, (ntri * 3, 3)).astype(np.float32)
        d.ops.append(("mesh", v.ravel(), np.arange(ntri * 3, dtype=np.uint32), None, None, 0))
        check_bvh(d)
