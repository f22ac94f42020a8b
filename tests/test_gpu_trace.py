"""GPU parity: LBVH + watertight traversal against the oracle's brute-force
intersector (hit primitive ids and t bit-exact; ties resolve identically by
the smaller primitive id)."""
import numpy as np
import pytest

from offlinerenderer_amd import scenes
from tests import gpu_util as G
from tests.scene_util import make_rays

pytestmark = pytest.mark.gpu


def _compare(t_g, p_g, t_o, p_o):
    assert np.array_equal(p_g, p_o), f"prim mismatch on {np.count_nonzero(p_g != p_o)} rays"
    assert np.array_equal(t_g.view(np.uint32), t_o.view(np.uint32)), "t mismatch"


@pytest.mark.parametrize("seed", [1, 2])
def test_cornell_closest_and_any_bit_exact(gpu, oracle, seed):
    d = scenes.cornell(64, 64, 1)
    ps = G.product_scene(d)
    os_ = oracle.build_oracle_scene(d, bvh=False)
    org4, dir4 = make_rays(d, 100_000, seed)
    bvh = ps.getBVH()
    t_g, p_g = G.trace_closest(bvh, org4, dir4)
    t_o, p_o = oracle.trace_closest(os_, org4, dir4, use_bvh=False)
    _compare(t_g, p_g, t_o, p_o)
    assert np.count_nonzero(p_o != 0xFFFFFFFF) > 10_000
    occ_g = G.trace_any(bvh, org4, dir4)
    occ_o = oracle.trace_any(os_, org4, dir4, use_bvh=False)
    assert np.array_equal(occ_g, occ_o)


def test_material_scene_closest_bit_exact(gpu, oracle):
    d = scenes.material_scene(96, 54, 1, tex_size=64)
    ps = G.product_scene(d)
    os_ = oracle.build_oracle_scene(d, bvh=True)
    org4, dir4 = make_rays(d, 60_000, 7)
    bvh = ps.getBVH()
    assert bvh.numTris == d.numTris
    t_g, p_g = G.trace_closest(bvh, org4, dir4)
    t_o, p_o = oracle.trace_closest(os_, org4, dir4, use_bvh=True)
    _compare(t_g, p_g, t_o, p_o)
    # brute force on a subset pins the oracle BVH too
    t_b, p_b = oracle.trace_closest(os_, org4[:3000], dir4[:3000], use_bvh=False)
    _compare(t_g[:3000], p_g[:3000], t_b, p_b)
    occ_g = G.trace_any(bvh, org4, dir4)
    occ_o = oracle.trace_any(os_, org4, dir4, use_bvh=True)
    assert np.array_equal(occ_g, occ_o)


def test_queue_indirection(gpu, oracle):
    d = scenes.cornell(8, 8, 1)
    ps = G.product_scene(d)
    os_ = oracle.build_oracle_scene(d, bvh=False)
    org4, dir4 = make_rays(d, 4096, 3)
    queue = np.random.default_rng(0).permutation(4096)[:1000].astype(np.uint32)
    t_g, p_g = G.trace_closest(ps.getBVH(), org4, dir4, queue=queue)
    t_o, p_o = oracle.trace_closest(os_, org4[queue], dir4[queue], use_bvh=False)
    _compare(t_g[queue], p_g[queue], t_o, p_o)


def test_large_incoherent_scene_trace(gpu, oracle):
    """200 displaced icospheres (1.02M triangles) in a 100^3 box: GPU LBVH +
    traversal against the oracle's BVH (itself pinned to brute force)."""
    d = scenes.large_scene(64, 36, 1, num_spheres=200, num_emitters=16)
    ps = G.product_scene(d)
    os_ = oracle.build_oracle_scene(d, bvh=True)
    bvh = ps.getBVH()
    assert bvh.numTris == d.numTris
    org4, dir4 = make_rays(d, 40_000, 5)
    t_g, p_g = G.trace_closest(bvh, org4, dir4)
    t_o, p_o = oracle.trace_closest(os_, org4, dir4, use_bvh=True)
    _compare(t_g, p_g, t_o, p_o)
    t_b, p_b = oracle.trace_closest(os_, org4[:300], dir4[:300], use_bvh=False)
    _compare(t_g[:300], p_g[:300], t_b, p_b)
    assert np.array_equal(G.trace_any(bvh, org4, dir4), oracle.trace_any(os_, org4, dir4, use_bvh=True))


def test_many_rays_per_lane_match_one_per_lane(gpu, oracle):
    """A launch of 9.4M rays refills lanes from each wave's chunk (several
    rays per lane on a 256-CU device); 1M-ray launches take one ray per lane.
    Both must give the same hits, and a slice is checked against the oracle."""
    d = scenes.material_scene(96, 54, 1, tex_size=64)
    ps = G.product_scene(d)
    bvh = ps.getBVH()
    n, m = 9 << 20, 1 << 20
    org4, dir4 = make_rays(d, n, 11)
    t_all, p_all = G.trace_closest(bvh, org4, dir4)
    occ_all = G.trace_any(bvh, org4, dir4)
    for k in range(0, n, m):
        t_k, p_k = G.trace_closest(bvh, org4[k:k + m], dir4[k:k + m])
        _compare(t_all[k:k + m], p_all[k:k + m], t_k, p_k)
        assert np.array_equal(occ_all[k:k + m], G.trace_any(bvh, org4[k:k + m], dir4[k:k + m]))
    os_ = oracle.build_oracle_scene(d, bvh=True)
    sl = slice(n - 20_000, n)
    t_o, p_o = oracle.trace_closest(os_, org4[sl], dir4[sl], use_bvh=True)
    _compare(t_all[sl], p_all[sl], t_o, p_o)
    assert np.array_equal(occ_all[sl], oracle.trace_any(os_, org4[sl], dir4[sl], use_bvh=True))


def test_rays_per_lane_cap_matches_one_per_lane(gpu):
    """A 28M-ray launch (three copies of 9.4M rays) reaches the cap of 12
    rays per lane on a 256-CU device; its hits equal those of 1M-ray launches
    (one ray per lane) bit for bit."""
    d = scenes.material_scene(96, 54, 1, tex_size=64)
    bvh = G.product_scene(d).getBVH()
    org1, dir1 = make_rays(d, 9 << 20, 13)
    org4, dir4 = np.concatenate([org1] * 3), np.concatenate([dir1] * 3)
    n, m = org4.shape[0], 1 << 20
    t_all, p_all = G.trace_closest(bvh, org4, dir4)
    for k in range(0, n, m):
        t_k, p_k = G.trace_closest(bvh, org4[k:k + m], dir4[k:k + m])
        _compare(t_all[k:k + m], p_all[k:k + m], t_k, p_k)
    assert np.count_nonzero(p_all != 0xFFFFFFFF) > n // 10


def test_closest_hit_categories(gpu):
    """wrk_trace_closest_categorized: hits equal wrk_trace_closest's, and the
    category at each queue position is the material kind of the hit
    triangle (unknown kinds -> emissive queue, bad ids -> finished), 0xFF on
    a miss."""
    import ctypes as C

    import torch

    import offlinerenderer_amd as wr
    from tests.test_gpu_bvh import d2h

    d = scenes.material_scene(96, 54, 1, tex_size=64)
    ps = G.product_scene(d)
    bvh = ps.getBVH()
    org4, dir4 = make_rays(d, 50_000, 13)
    t_ref, p_ref = G.trace_closest(bvh, org4, dir4)
    n = org4.shape[0]
    rng = np.random.default_rng(1)
    queue = rng.permutation(n).astype(np.uint32)
    o, dd = G.dev(org4), G.dev(dir4)
    q = G.dev(queue)
    hits = torch.zeros((n, 2), dtype=torch.int32, device="cuda")
    cats = torch.full((n,), 0x77, dtype=torch.uint8, device="cuda")
    G.check(wr.lib().wrk_trace_closest_categorized(C.byref(bvh), G.vp(o), G.vp(dd), G.vp(q), n, None,
                                                   G.vp(hits), G.vp(cats), None, G.stream()))
    torch.cuda.synchronize()
    h = hits.cpu().numpy().view(np.uint32)
    assert np.array_equal(h[:, 1], p_ref) and np.array_equal(h[:, 0].view(np.float32).view(np.uint32),
                                                             t_ref.view(np.uint32))
    view = ps.getDeviceView()
    mats = d2h(view.materials, view.numMaterials * 64).view(np.uint32).reshape(-1, 16)
    kind = mats[:, 9]  # wrk_material.type
    tri_mat = d2h(view.triMaterialIds, view.numTris * 4).view(np.uint32)
    prim = p_ref[queue]
    hit = prim != 0xFFFFFFFF
    m = tri_mat[prim[hit]]
    want = np.where(m >= view.numMaterials, 5, np.where(kind[np.minimum(m, view.numMaterials - 1)] <= 4,
                                                        kind[np.minimum(m, view.numMaterials - 1)], 1))
    got = cats.cpu().numpy()
    assert np.array_equal(got[hit], want)
    assert np.all(got[~hit] == 0xFF)
    assert len(np.unique(want)) >= 3


@pytest.mark.parametrize("tagged", [True, False])
def test_resolving_traversals_match_trace_then_classify(gpu, tagged):
    """wrk_trace_any_resolve + wrk_trace_closest_resolve leave the path pool
    (radiance, flushed results, hit records of hits) and the categories
    exactly as wrk_trace_any + wrk_trace_closest_categorized + wrk_classify
    do, for a random pool state: trace-only, trace + shadow and shadow-only
    entries, occluded and unoccluded shadow rays, misses, firefly clamps per
    slot, NEE-done flags; with tagged triangle records and without."""
    import ctypes as C

    import torch

    import offlinerenderer_amd as wr

    d = scenes.material_scene(96, 54, 1, tex_size=64)
    ps = G.product_scene(d)
    bvh = ps.getBVH()
    view = ps.getDeviceView()
    cap = 300_000
    rng = np.random.default_rng(7 + tagged)
    org4, dir4 = make_rays(d, cap, 21)
    sorg, sdir = make_rays(d, cap, 22)
    sdir[:, 3] = np.where(rng.uniform(size=cap) < 0.5, sdir[:, 3], rng.uniform(0.05, 20.0, cap))
    f = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()  # noqa: E731
    state = np.zeros((cap, 4), np.uint32)
    state[:, 0] = rng.integers(0, 1 << 32, cap, dtype=np.uint64)
    slots = rng.integers(0, 4, cap).astype(np.uint32)
    nee = (rng.uniform(size=cap) < 0.3).astype(np.uint32) << 18
    state[:, 2] = (slots << 24) | nee | rng.integers(0, 8, cap).astype(np.uint32)
    thr, rad, con = (rng.uniform(0, 3, (cap, 4)).astype(np.float32) for _ in range(3))
    prm = np.zeros((256, 8), np.uint32)
    prm.view(np.float32)[:4, 5] = [0.0, 0.5, 2.0, 0.0]  # fireflyClamp per slot
    n = 240_000
    nT, nTS, nS = 120_000, 70_000, 50_000
    queue = rng.permutation(cap)[:n].astype(np.uint32)

    def pool():
        bufs = dict(org=f(org4), dir=f(dir4), thr=f(thr), rad=f(rad), state=G.dev(state.view(np.int32)),
                    contrib=f(con), savedOrg=f(sorg), savedDir=f(sdir),
                    hit=torch.full((cap, 2), -1, dtype=torch.int32, device="cuda"),
                    occluded=torch.zeros(cap, dtype=torch.uint8, device="cuda"),
                    result=torch.zeros((cap, 4), dtype=torch.float32, device="cuda"))
        P = wr.wrk_paths()
        for k, t in bufs.items():
            setattr(P, k, t.data_ptr())
        P.capacity = cap
        return P, bufs

    b = wr.wrk_bvh.from_buffer_copy(bvh)
    if not tagged:
        # untagged triangle records: the category comes from the material table
        raw = torch.zeros(b.numTris * 12, dtype=torch.float32, device="cuda")
        import tests.test_gpu_bvh as TB
        recs = TB.d2h(b.tris, b.numTris * 48).view(np.float32).reshape(b.numTris, 3, 4).copy()
        recs[:, 1, 3] = 0.0
        raw = torch.from_numpy(recs.ravel()).cuda()
        b.tris = raw.data_ptr()
    prm_d = G.dev(prm.view(np.int32))
    q = G.dev(queue)
    L = wr.lib()

    P1, B1 = pool()
    cat1 = torch.full((n,), 0x77, dtype=torch.uint8, device="cuda")
    G.check(L.wrk_trace_any(C.byref(b), C.c_void_p(P1.savedOrg), C.c_void_p(P1.savedDir), G.vp(q[nT:]), nTS + nS,
                            None, C.c_void_p(P1.occluded), G.stream()))
    G.check(L.wrk_trace_closest_categorized(C.byref(b), C.c_void_p(P1.org), C.c_void_p(P1.dir), G.vp(q), nT + nTS,
                                            None, C.c_void_p(P1.hit), G.vp(cat1), None, G.stream()))
    G.check(L.wrk_classify(C.byref(view), G.vp(prm_d), C.byref(P1), G.vp(q), nT, nTS, nS, G.vp(cat1), 1,
                           G.stream()))
    P2, B2 = pool()
    cat2 = torch.full((n,), 0x77, dtype=torch.uint8, device="cuda")
    G.check(L.wrk_trace_any_resolve(C.byref(b), G.vp(prm_d), C.byref(P2), G.vp(q[nT:]), nTS, nS,
                                    C.c_void_p(cat2.data_ptr() + nT), None, G.stream()))
    G.check(L.wrk_trace_closest_resolve(C.byref(b), C.byref(view), G.vp(prm_d), C.byref(P2), G.vp(q), nT + nTS,
                                        G.vp(cat2), None, G.stream()))
    torch.cuda.synchronize()
    c1, c2 = cat1.cpu().numpy(), cat2.cpu().numpy()
    assert np.array_equal(c1, c2)
    for k in ("rad", "result"):
        assert np.array_equal(B1[k].cpu().numpy().view(np.uint32), B2[k].cpu().numpy().view(np.uint32)), k
    hit1, hit2 = B1["hit"].cpu().numpy().view(np.uint32), B2["hit"].cpu().numpy().view(np.uint32)
    hits = c1 < 5
    hq = queue[:nT + nTS][hits[:nT + nTS]]
    assert np.array_equal(hit1[hq], hit2[hq])
    # the mix is exercised: misses, hits of several kinds, resumed and flushed shadow-only entries
    assert (c1[:nT + nTS] == 5).sum() > 1000 and len(np.unique(c1[:nT + nTS])) >= 4
    assert (c1[nT + nTS:] == 0).sum() > 1000 and (c1[nT + nTS:] == 5).sum() > 1000
