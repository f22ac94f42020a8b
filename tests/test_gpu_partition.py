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
def test_slot_finish_matches_histogram(gpu, n):
    rng = np.random.default_rng(1234 + n)
    cap = max(2 * n, 64)
    # path states: slot in bits 24..31 of state.z; runs of one slot as the
    # partition leaves them, then scattered slots
    slot_of = rng.integers(0, 256, size=cap).astype(np.uint32)
    slot_of[: cap // 2] = np.sort(slot_of[: cap // 2])
    state = np.zeros((cap, 4), np.uint32)
    state[:, 2] = (slot_of << 24) | rng.integers(0, 1 << 24, size=cap).astype(np.uint32)
    finished = rng.permutation(cap)[:n].astype(np.uint32)
    finished[: n // 2] = np.sort(finished[: n // 2])
    m = max(n - 5, 0)  # device-side count shorter than the launch size
    live0 = rng.integers(1 << 20, 1 << 24, size=256).astype(np.uint32)
    st = G.dev(state)
    q = G.dev(finished if n else np.zeros(1, np.uint32))
    cnt = G.dev(np.array([m], np.uint32))
    live = G.dev(live0)
    paths = wr.wrk_paths()
    paths.state = st.data_ptr()
    paths.capacity = cap
    G.check(wr.lib().wrk_slot_finish(G.vp(q), n, G.vp(cnt), C.byref(paths), G.vp(live), G.stream()))
    torch.cuda.synchronize()
    got = live.cpu().numpy().view(np.uint32)
    want = live0 - np.bincount(slot_of[finished[:m]], minlength=256).astype(np.uint32)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("n", [0, 1, 1000, 3000017])
def test_slot_finish_map_matches_histogram(gpu, n):
    """wrk_slot_finish_map: the slot of pool entry p is slotMap[p >> 8];
    entries past the map are skipped (a few are planted), and more entries
    than the map covers are refused."""
    rng = np.random.default_rng(4321 + n)
    granules = max((2 * n + 255) // 256, 1)
    slot_map = rng.integers(0, 256, size=granules).astype(np.uint8)
    slot_map[: granules // 2] = np.sort(slot_map[: granules // 2])
    finished = rng.permutation(granules * 256)[:n].astype(np.uint32)
    finished[: n // 2] = np.sort(finished[: n // 2])
    if n > 10:
        finished[[3, n // 2, n - 7]] = [granules * 256, granules * 256 + 999, 0xFFFFFF00]
    m = max(n - 5, 0)  # device-side count shorter than the launch size
    live0 = rng.integers(1 << 20, 1 << 24, size=256).astype(np.uint32)
    q = G.dev(finished if n else np.zeros(1, np.uint32))
    cnt = G.dev(np.array([m], np.uint32))
    live = G.dev(live0)
    smap = G.dev(slot_map)
    G.check(wr.lib().wrk_slot_finish_map(G.vp(q), n, G.vp(cnt), G.vp(smap), granules, G.vp(live), G.stream()))
    torch.cuda.synchronize()
    got = live.cpu().numpy().view(np.uint32)
    inside = finished[:m][(finished[:m] >> 8) < granules]
    want = live0 - np.bincount(slot_map[inside >> 8], minlength=256).astype(np.uint32)
    assert np.array_equal(got, want)
    if n:
        rc = wr.lib().wrk_slot_finish_map(G.vp(q), 256 * granules + 1, G.vp(cnt), G.vp(smap), granules, G.vp(live),
                                          G.stream())
        assert rc != 0 and b"slot map" in wr.lib().wrk_last_error()



@pytest.mark.parametrize("K", [6, 13])
@pytest.mark.parametrize("off", [0, 1, 2, 3])
@pytest.mark.parametrize("n", [1023, 1027, 5000, 777_777])
def test_partition_production_widths_at_queue_offsets(gpu, K, off, n):
    """The engine partitions each scene's part of the queue with K = 6 from
    `category + a`, `queue + a` (any a), and the whole queue with K = 13.
    A category pointer that is not 4-byte aligned takes the bytewise kernels;
    an aligned one the 4-per-lane kernels. Both must give numpy's stable
    partition, with n around tile (1024) and word edges and over several
    scan segments."""
    rng = np.random.default_rng(K * 7919 + off * 131 + n)
    cat = rng.integers(0, K + 1, size=n).astype(np.uint8)   # category K (and above) is dropped
    cat[rng.uniform(size=n) < 0.3] = K - 1                    # a heavy last category, like "finished"
    items = rng.permutation(n).astype(np.uint32)
    L = wr.lib()
    # buffers padded on both sides: the kernels see only [off, off + n)
    cat_big = np.full(n + 8, 0xEE, np.uint8)
    cat_big[off:off + n] = cat
    items_big = np.full(n + 8, 0xDEADBEEF, np.uint32)
    items_big[off:off + n] = items
    cat_d, items_d = G.dev(cat_big), G.dev(items_big)
    out_d = torch.full((n + 8,), -1, dtype=torch.int32, device="cuda")
    offs = torch.zeros(K + 1, dtype=torch.int32, device="cuda")
    sb = L.wrk_partition_scratch_bytes(n, K)
    scratch = torch.empty(sb, dtype=torch.uint8, device="cuda")
    import ctypes as C

    G.check(L.wrk_partition(C.c_void_p(cat_d.data_ptr() + off), C.c_void_p(items_d.data_ptr() + 4 * off), n, K,
                            C.c_void_p(out_d.data_ptr() + 4 * off), G.vp(offs), G.vp(scratch), sb, G.stream()))
    torch.cuda.synchronize()
    o = offs.cpu().numpy()
    got = out_d.cpu().numpy().view(np.uint32)
    expect_offs = np.concatenate([[0], np.cumsum([np.count_nonzero(cat == k) for k in range(K)])])
    assert np.array_equal(o, expect_offs)
    for k in range(K):
        assert np.array_equal(got[off + o[k]:off + o[k + 1]], items[cat == k]), k
    # nothing written outside the partitioned range
    assert np.all(got[:off] == 0xFFFFFFFF) and np.all(got[off + o[K]:] == 0xFFFFFFFF)
