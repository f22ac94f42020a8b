"""wrk_partition_noalloc (the allocation-free variant of wrk_partition)
against wrk_partition with a caller-supplied scratch buffer, through the C
ABI: two calls in a row on one GPU, the second with a smaller input, equal
the scratch-buffer entry point's queues and offsets bit for bit; sizes past
the maximum are rejected with a clear error."""
import ctypes as C

import numpy as np
import pytest
import torch

import offlinerenderer_amd as wr
from tests import gpu_util as G

pytestmark = pytest.mark.gpu


def _inputs(n, K, seed, off):
    rng = np.random.default_rng(seed)
    cat = rng.integers(0, K + 2, size=n).astype(np.uint8)
    items = rng.permutation(n).astype(np.uint32)
    cat_big = np.zeros(n + 8, np.uint8)
    cat_big[off:off + n] = cat
    return cat, items, G.dev(cat_big), G.dev(items)


def _run(fn_noalloc, cat_d, off, items_d, n, K):
    L = wr.lib()
    out = torch.full((max(n, 1),), -1, dtype=torch.int32, device="cuda")
    offs = torch.full((K + 1,), -1, dtype=torch.int32, device="cuda")
    cptr = C.c_void_p(cat_d.data_ptr() + off)
    if fn_noalloc:
        G.check(L.wrk_partition_noalloc(cptr, G.vp(items_d), n, K, G.vp(out), G.vp(offs), G.stream()))
    else:
        sb = L.wrk_partition_scratch_bytes(n, K)
        scratch = torch.empty(sb, dtype=torch.uint8, device="cuda")
        G.check(L.wrk_partition(cptr, G.vp(items_d), n, K, G.vp(out), G.vp(offs), G.vp(scratch), sb, G.stream()))
    torch.cuda.synchronize()
    return out.cpu().numpy().view(np.uint32)[:n].copy(), offs.cpu().numpy().copy()


def test_noalloc_partition_equals_scratch_partition_twice(gpu):
    L = wr.lib()
    assert L.wrk_partition_noalloc_max_items() == 1 << 26
    # first call large (several scan segments, K = 13), second smaller
    # (single-segment scan, K = 6, an unaligned category pointer)
    for n, K, off, seed in ((3_000_017, 13, 0, 1), (5000, 6, 1, 2), (0, 4, 0, 3)):
        cat, items, cat_d, items_d = _inputs(n, K, seed, off)
        a_out, a_offs = _run(True, cat_d, off, items_d, n, K)
        b_out, b_offs = _run(False, cat_d, off, items_d, n, K)
        assert np.array_equal(a_offs, b_offs) and np.array_equal(a_out, b_out)
        expect = np.concatenate([[0], np.cumsum([np.count_nonzero(cat == k) for k in range(K)])])
        assert np.array_equal(a_offs, expect)
        for k in range(K):
            assert np.array_equal(a_out[a_offs[k]:a_offs[k + 1]], items[cat == k])


def test_noalloc_partition_rejects_oversize(gpu):
    L = wr.lib()
    dummy = torch.zeros(16, dtype=torch.int32, device="cuda")
    rc = L.wrk_partition_noalloc(G.vp(dummy), None, (1 << 26) + 1, 6, G.vp(dummy), G.vp(dummy), G.stream())
    assert rc != 0 and b"exceeds the maximum" in L.wrk_last_error()
    rc = L.wrk_partition_noalloc(G.vp(dummy), None, 16, 17, G.vp(dummy), G.vp(dummy), G.stream())
    assert rc != 0 and b"numCategories" in L.wrk_last_error()
