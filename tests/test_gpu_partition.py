"""Ordered compaction / K-way partition kernels against the oracle's
copy_if and numpy stable partition (bit-exact queues), including empty,
ragged and multi-segment sizes."""
import ctypes as C

import numpy as np
import pytest
import torch

import offlinerenderer_amd as wr
from tests import gpu_util as G

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [0, 1, 1023, 1024, 1025, 77777, 9_000_001])
def test_compact_matches_copy_if(gpu, oracle, n):
    rng = np.random.default_rng(n)
    keep = (rng.uniform(size=n) < 0.37).astype(np.uint8)
    items = rng.integers(0, 1 << 31, size=n, dtype=np.uint32)
    L = wr.lib()
    sb = L.wrk_compact_scratch_bytes(n)
    scratch = torch.empty(max(sb, 1), dtype=torch.uint8, device="cuda")
    out = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int32, device="cuda")
    kd = G.dev(keep) if n else torch.zeros(1, dtype=torch.uint8, device="cuda")
    it = G.dev(items) if n else torch.zeros(1, dtype=torch.int32, device="cuda")
    G.check(L.wrk_compact(G.vp(kd), G.vp(it), n, G.vp(out), G.vp(cnt), G.vp(scratch), sb, G.stream()))
    torch.cuda.synchronize()
    c = int(cnt.item())
    ref = oracle.compact(keep, items) if n else np.zeros(0, np.uint32)
    assert c == ref.size
    assert np.array_equal(out.cpu().numpy().view(np.uint32)[:c], ref)


@pytest.mark.parametrize("K", [2, 5, 8])
@pytest.mark.parametrize("n", [5000, 3_000_017])
def test_partition_is_stable(gpu, K, n):
    rng = np.random.default_rng(K * 1000 + n % 997)
    cat = rng.integers(0, K + 2, size=n).astype(np.uint8)   # categories >= K are dropped
    items = rng.permutation(n).astype(np.uint32)
    L = wr.lib()
    sb = L.wrk_partition_scratch_bytes(n, K)
    scratch = torch.empty(sb, dtype=torch.uint8, device="cuda")
    out = torch.zeros(n, dtype=torch.int32, device="cuda")
    offs = torch.zeros(K + 1, dtype=torch.int32, device="cuda")
    cat_d, items_d = G.dev(cat), G.dev(items)  # keep the device buffers alive across the launch
    G.check(L.wrk_partition(G.vp(cat_d), G.vp(items_d), n, K, G.vp(out), G.vp(offs), G.vp(scratch), sb, G.stream()))
    torch.cuda.synchronize()
    o = offs.cpu().numpy()
    got = out.cpu().numpy().view(np.uint32)
    expect_offs = np.concatenate([[0], np.cumsum([np.count_nonzero(cat == k) for k in range(K)])])
    assert np.array_equal(o, expect_offs)
    for k in range(K):
        assert np.array_equal(got[o[k]:o[k + 1]], items[cat == k])


@pytest.mark.parametrize("n", [0, 1, 1000, 3000017])
def test_slot_count_matches_histogram(gpu, n):
    rng = np.random.default_rng(1234 + n)
    cap = max(2 * n, 64)
    nr = 37
    firsts = np.concatenate([[0], np.sort(rng.choice(np.arange(1, cap), nr - 1, replace=False))]).astype(np.uint32)
    slots = rng.permutation(256)[:nr].astype(np.uint32)
    queue = rng.permutation(cap)[:n].astype(np.uint32)
    queue[: n // 2] = np.sort(queue[: n // 2])  # long runs of one range, then scattered entries
    m = max(n - 5, 0)  # device-side count shorter than the launch size
    q = G.dev(queue if n else np.zeros(1, np.uint32))
    ranges = G.dev(np.concatenate([firsts, slots]))
    cnt = G.dev(np.array([m], np.uint32))
    counts = G.dev(np.full(256, 7, np.uint32))
    G.check(wr.lib().wrk_slot_count(G.vp(q), n, G.vp(cnt), G.vp(ranges), nr, G.vp(counts), G.stream()))
    torch.cuda.synchronize()
    got = counts.cpu().numpy().view(np.uint32)
    r = np.searchsorted(firsts, queue[:m], side="right") - 1
    want = np.bincount(slots[r], minlength=256).astype(np.uint32)
    assert np.array_equal(got, want)
