"""The BVH build's building blocks through the C ABI: the LDS-staged radix
sort against numpy's stable argsort (std::stable_sort order: equal keys keep
their input order) and the exclusive scan against numpy's cumulative sum,
including empty, single, tile-edge and 10M-entry inputs. Bit-exact."""
import ctypes as C

import numpy as np
import pytest
import torch

import offlinerenderer_amd as wr
from tests import gpu_util as G

pytestmark = pytest.mark.gpu


def _sort(keys, vals, begin, end):
    L = wr.lib()
    n = keys.size
    k = G.dev(keys.view(np.int64)) if n else torch.zeros(1, dtype=torch.int64, device="cuda")
    v = G.dev(vals.view(np.int32)) if n else torch.zeros(1, dtype=torch.int32, device="cuda")
    kt, vt = torch.empty_like(k), torch.empty_like(v)
    sb = L.wrk_radix_sort_scratch_bytes(n)
    scratch = torch.empty(max(sb, 1), dtype=torch.uint8, device="cuda")
    G.check(L.wrk_radix_sort_pairs_u64(G.vp(k), G.vp(v), G.vp(kt), G.vp(vt), n, begin, end, G.vp(scratch), sb,
                                       G.stream()))
    torch.cuda.synchronize()
    return k.cpu().numpy().view(np.uint64)[:n], v.cpu().numpy().view(np.uint32)[:n]


@pytest.mark.parametrize("n", [0, 1, 2, 1000, 2048, 2049, 100_003, 10_240_064])
@pytest.mark.parametrize("begin,end", [(0, 63), (0, 17), (5, 40)])
def test_radix_sort_matches_stable_sort(gpu, n, begin, end):
    rng = np.random.default_rng(n * 7 + begin)
    # a mix of distinct 63-bit codes and heavy duplicates (equal Morton codes)
    keys = rng.integers(0, 1 << 63, size=n, dtype=np.uint64)
    dup = rng.uniform(size=n) < 0.3
    keys[dup] = rng.integers(0, 64, size=int(dup.sum()), dtype=np.uint64) << np.uint64(begin)
    vals = rng.permutation(n).astype(np.uint32)
    ks, vs = _sort(keys, vals, begin, end)
    sel = (keys >> np.uint64(begin)) & np.uint64((1 << (end - begin)) - 1)
    order = np.argsort(sel, kind="stable")
    assert np.array_equal(vs, vals[order])
    assert np.array_equal(ks, keys[order])


def test_radix_sort_rejects_bad_arguments(gpu):
    L = wr.lib()
    d = torch.zeros(64, dtype=torch.int64, device="cuda")
    assert L.wrk_radix_sort_pairs_u64(G.vp(d), G.vp(d), G.vp(d), G.vp(d), 8, 10, 70, G.vp(d), 512, G.stream()) != 0
    assert L.wrk_radix_sort_pairs_u64(G.vp(d), G.vp(d), G.vp(d), G.vp(d), 8, 0, 63, G.vp(d), 1, G.stream()) != 0
    assert b"scratch" in L.wrk_last_error()


@pytest.mark.parametrize("n", [0, 1, 8191, 8192, 8193, 1_000_001, 20_000_000])
def test_exclusive_scan_matches_cumsum(gpu, n):
    L = wr.lib()
    rng = np.random.default_rng(n)
    a = rng.integers(0, 3, size=n).astype(np.uint32)
    x = G.dev(a.view(np.int32)) if n else torch.zeros(1, dtype=torch.int32, device="cuda")
    out = torch.full((max(n, 1),), -1, dtype=torch.int32, device="cuda")
    tot = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    sb = L.wrk_exclusive_scan_scratch_bytes(n)
    scratch = torch.empty(max(sb, 4), dtype=torch.uint8, device="cuda")
    G.check(L.wrk_exclusive_scan_u32(G.vp(x), G.vp(out), n, G.vp(tot), G.vp(scratch), sb, G.stream()))
    torch.cuda.synchronize()
    got = out.cpu().numpy().view(np.uint32)[:n]
    ref = (np.cumsum(a, dtype=np.uint64) - a).astype(np.uint32)
    assert np.array_equal(got, ref)
    if n:
        assert int(tot.cpu().numpy().view(np.uint32)[0]) == int(a.sum())
    # in place
    if n:
        G.check(L.wrk_exclusive_scan_u32(G.vp(x), G.vp(x), n, None, G.vp(scratch), sb, G.stream()))
        torch.cuda.synchronize()
        assert np.array_equal(x.cpu().numpy().view(np.uint32), ref)
