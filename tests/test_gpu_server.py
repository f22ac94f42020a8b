"""Continuous batching through wr::RenderServer: a request mixed with
requests of other scenes, resolutions and sample counts, admitted at a later
iteration, or split into chunks, yields bit for bit the image it yields
rendered alone with wr::Renderer (same per-path PCG seeds, accumulation in
sample order per pixel)."""
import numpy as np
import pytest

import offlinerenderer_amd as wr
from offlinerenderer_amd import scenes
from tests import gpu_util as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def world(gpu):
    descs = {
        "cornell": scenes.cornell(64, 64, 4),
        "material": scenes.material_scene(96, 54, 2, tex_size=64),
        "large": scenes.large_scene(64, 36, 1, num_spheres=200, num_emitters=16),
    }
    return {k: (d, G.product_scene(d)) for k, d in descs.items()}


def _requests(world):
    dc, sc = world["cornell"]
    dm, sm = world["material"]
    dl, sl = world["large"]
    moved = dict(dc.camera)
    moved["position"] = (0.3, 0.2, 2.6)
    return [
        (sc, wr.Camera(**dc.camera), G.params(dc)),
        (sm, wr.Camera(**dm.camera), G.params(dm)),
        (sc, wr.Camera(**moved), wr.RenderParams(width=32, height=48, spp=3, maxBounces=8)),
        (sl, wr.Camera(**dl.camera), G.params(dl)),
        (sm, wr.Camera(**dm.camera), wr.RenderParams(width=40, height=30, spp=1, maxBounces=4)),
    ]


def _alone(reqs, cap):
    r = wr.Renderer()
    r.setWavefrontConfig(wr.WavefrontConfig(iterationCap=cap, maxPathsInFlight=1 << 20))
    return [r.render(s, c, p) for s, c, p in reqs]


def _serve(reqs, cfg, stagger):
    srv = wr.RenderServer(cfg)
    ids, out = [], {}
    pending = list(reqs)
    steps = 0
    while pending or len(out) < len(ids):
        # later requests arrive while earlier ones are in flight
        take = 1 if stagger else len(pending)
        for s, c, p in pending[:take]:
            ids.append(srv.submit(s, c, p))
        pending = pending[take:]
        srv.step()
        steps += 1
        done = srv.takeCompleted()
        assert all(srv.status(rid)["done"] for rid in done)
        for rid in done:
            out[rid] = srv.fetch(rid)
        assert steps < 10000
    return [out[rid] for rid in ids], srv


def _same(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("shortest_first,round_robin", [(False, False), (True, False), (False, True), (True, True)])
def test_mixed_requests_match_alone(world, shortest_first, round_robin):
    reqs = _requests(world)
    ref = _alone(reqs, cap=0)
    got, srv = _serve(reqs, wr.ServerConfig(maxPathsInFlight=1 << 20, shortestFirst=shortest_first,
                                            roundRobin=round_robin), stagger=True)
    for i, (a, b) in enumerate(zip(got, ref)):
        assert _same(a, b), f"request {i} differs from its stand-alone render"
    assert srv.numPending() == 0 and srv.numInFlight() == 0
    st = srv.getStats()
    assert st["closestRays"] > 0 and st["iterations"] > 0


def test_chunked_small_pool_matches_alone(world):
    reqs = _requests(world)
    ref = _alone(reqs, cap=0)
    # pool smaller than the largest request; chunks smaller than the pool
    for rr in (False, True):
        got, _ = _serve(reqs, wr.ServerConfig(maxPathsInFlight=12000, maxChunkPaths=5000, roundRobin=rr),
                        stagger=False)
        for i, (a, b) in enumerate(zip(got, ref)):
            assert _same(a, b), f"request {i} differs when split into chunks (round robin {rr})"


def test_in_order_admission(world):
    """Without round-robin, a request is first admitted only after every
    request submitted before it has had all its chunks admitted."""
    reqs = _requests(world)
    ref = _alone(reqs, cap=0)
    srv = wr.RenderServer(wr.ServerConfig(maxPathsInFlight=12000, maxChunkPaths=3000))
    ids = [srv.submit(s, c, p) for s, c, p in reqs]
    admitted_before = []  # per step: which requests have been admitted so far
    for _ in range(10000):
        more = srv.step()
        admitted_before.append([srv.status(i)["admitted"] for i in ids])
        if not more:
            break
    for row in admitted_before:
        # admitted requests form a prefix of the submission order
        k = row.index(False) if False in row else len(row)
        assert all(not a for a in row[k:])
    for i, rid in enumerate(ids):
        assert srv.status(rid)["done"]
        assert _same(srv.fetch(rid), ref[i])


def test_backfill_budget(world):
    """With a backfill budget, the requests behind the oldest unfinished one
    hold at most that many pool entries while it has paths in flight; every
    image still equals its stand-alone render."""
    reqs = _requests(world)
    ref = _alone(reqs, cap=0)
    budget = 2048
    srv = wr.RenderServer(wr.ServerConfig(maxPathsInFlight=1 << 16, maxChunkPaths=8192, backfillPaths=budget))
    ids = [srv.submit(s, c, p) for s, c, p in reqs]
    out = {}
    for _ in range(10000):
        more = srv.step()
        for rid in srv.takeCompleted():
            out[rid] = srv.fetch(rid)
        if not more:
            break
    assert sorted(out) == sorted(ids)
    for i, rid in enumerate(ids):
        assert _same(out[rid], ref[i]), f"request {i} differs with a backfill budget"
    for stagger in (False, True):
        got, _ = _serve(reqs, wr.ServerConfig(maxPathsInFlight=1 << 16, maxChunkPaths=8192, backfillPaths=budget,
                                              fusedMaxPaths=4096), stagger=stagger)
        for i, (a, b) in enumerate(zip(got, ref)):
            assert _same(a, b), f"request {i} differs with a backfill budget and fused requests (stagger {stagger})"


def test_reference_mode_cap_with_staggered_admission(world):
    reqs = _requests(world)[:3]
    ref = _alone(reqs, cap=8)
    got, _ = _serve(reqs, wr.ServerConfig(maxPathsInFlight=1 << 20, iterationCap=8), stagger=True)
    for i, (a, b) in enumerate(zip(got, ref)):
        assert _same(a, b), f"request {i} differs under the 8-iteration cap"


def test_memory_budget_back_pressure(world):
    reqs = _requests(world)[:3]
    ref = _alone(reqs, cap=0)
    # accumulation + image buffers take 24 B per pixel: room for one request
    budget = max(p.width * p.height for _, _, p in reqs) * 24
    got, srv = _serve(reqs, wr.ServerConfig(maxPathsInFlight=1 << 20, memoryBudgetBytes=budget), stagger=False)
    for i, (a, b) in enumerate(zip(got, ref)):
        assert _same(a, b)


def test_server_rejects_bad_requests(world):
    srv = wr.RenderServer(wr.ServerConfig(maxPathsInFlight=4096))
    dc, sc = world["cornell"]
    with pytest.raises(RuntimeError):
        srv.submit(sc, wr.Camera(**dc.camera), wr.RenderParams(width=0, height=4, spp=1))
    assert not srv.step()


def test_renderer_small_chunks_match_one_batch(world):
    dm, sm = world["material"]
    cam, prm = wr.Camera(**dm.camera), G.params(dm)
    one = wr.Renderer()
    one.setWavefrontConfig(wr.WavefrontConfig(iterationCap=0, maxPathsInFlight=1 << 20))
    a = one.render(sm, cam, prm)
    # several chunks in flight at once; a chunk that finishes before its
    # predecessor waits so the per-pixel sum keeps the sample order
    many = wr.Renderer()
    many.setWavefrontConfig(wr.WavefrontConfig(iterationCap=0, maxPathsInFlight=6000, maxChunkPaths=1500))
    b = many.render(sm, cam, prm)
    assert _same(a, b)
    assert many.getStats()["batches"] >= (dm.width * dm.height * dm.spp) // 1500


@pytest.mark.parametrize("stagger,overlap", [(False, False), (True, False), (False, True), (True, True)])
def test_express_lane_matches_alone(world, stagger, overlap):
    """Requests of at most expressMaxPaths paths run in the express pool
    (here split into chunks, a few iterations per step, before or during
    the main iteration) and still equal their stand-alone renders."""
    reqs = _requests(world)
    ref = _alone(reqs, cap=0)
    cfg = wr.ServerConfig(maxPathsInFlight=1 << 20, expressMaxPaths=5000, expressPathsInFlight=4096,
                          expressIterations=3, expressOverlap=overlap)
    got, srv = _serve(reqs, cfg, stagger=stagger)
    for i, (a, b) in enumerate(zip(got, ref)):
        assert _same(a, b), f"request {i} differs (express lane, stagger {stagger}, overlap {overlap})"
    assert srv.numPending() == 0 and srv.numInFlight() == 0
    st = srv.getStats()
    assert st["closestRays"] > 0 and st["iterations"] > 0


def test_express_request_passes_long_one(world):
    """A short request submitted behind a long one that fills the pool
    finishes first when it takes the express lane, and after the long one
    without it (in-order admission)."""
    dc, sc = world["cornell"]
    dm, sm = world["material"]
    long_req = (sc, wr.Camera(**dc.camera), wr.RenderParams(width=64, height=64, spp=16, maxBounces=8))
    short_req = (sm, wr.Camera(**dm.camera), wr.RenderParams(width=32, height=32, spp=1, maxBounces=8))
    ref = _alone([long_req, short_req], cap=0)
    order = {}
    for express in (0, 2048):
        srv = wr.RenderServer(wr.ServerConfig(maxPathsInFlight=1 << 16, expressMaxPaths=express))
        ids = [srv.submit(*long_req), srv.submit(*short_req)]
        done, imgs = [], {}
        for _ in range(10000):
            more = srv.step()
            for rid in srv.takeCompleted():
                done.append(rid)
                imgs[rid] = srv.fetch(rid)
            if not more:
                break
        assert len(done) == 2
        order[express] = [ids.index(r) for r in done]
        for k, rid in enumerate(ids):
            assert _same(imgs[rid], ref[k])
    assert order[2048] == [1, 0]
    assert order[0] == [0, 1]


def test_express_lane_shares_memory_budget(world):
    """Both lanes charge one budget: with room for one request's buffers, a
    request is never admitted while another holds its buffers."""
    dc, sc = world["cornell"]
    dm, sm = world["material"]
    reqs = [(sc, wr.Camera(**dc.camera), wr.RenderParams(width=64, height=64, spp=2, maxBounces=8)),
            (sm, wr.Camera(**dm.camera), wr.RenderParams(width=32, height=48, spp=1, maxBounces=8)),
            (sc, wr.Camera(**dc.camera), wr.RenderParams(width=64, height=64, spp=1, maxBounces=8))]
    ref = _alone(reqs, cap=0)
    budget = 64 * 64 * 24
    srv = wr.RenderServer(wr.ServerConfig(maxPathsInFlight=1 << 16, memoryBudgetBytes=budget,
                                          expressMaxPaths=2000))
    ids = [srv.submit(*r) for r in reqs]
    fetched = {}
    for _ in range(10000):
        more = srv.step()
        holding = [rid for rid in ids if rid not in fetched and srv.status(rid)["admitted"]]
        assert len(holding) <= 1, "two requests hold buffers beyond the budget"
        for rid in srv.takeCompleted():
            fetched[rid] = srv.fetch(rid)
        if not more:
            break
    for k, rid in enumerate(ids):
        assert _same(fetched[rid], ref[k])


def test_mixed_requests_match_oracle_images(world, oracle):
    """A served image equals the oracle's image of the same request bit for
    bit: samples summed per pixel in sample order (renderer.cpp:190) and
    scaled by 1/spp (renderer.cpp:308-313). The server runs the requests
    mixed, staggered and split into chunks."""
    reqs = _requests(world)
    got, _ = _serve(reqs, wr.ServerConfig(maxPathsInFlight=12000, maxChunkPaths=5000), stagger=True)
    dc, _ = world["cornell"]
    dm, _ = world["material"]
    moved = dict(dc.camera)
    moved["position"] = (0.3, 0.2, 2.6)
    checks = [(0, dc, dc.camera, 64, 64, 4, 8), (2, dc, moved, 32, 48, 3, 8), (4, dm, dm.camera, 40, 30, 1, 4)]
    for i, desc, cam, w, h, spp, depth in checks:
        os_ = oracle.build_oracle_scene(desc, bvh=True)
        prm = oracle.params_in(w, h, depth, 3.0, iterationCap=0)
        ref = oracle.render_image(os_, cam, prm, spp) * np.float32(1.0 / spp)
        assert _same(got[i], ref.astype(np.float32)), f"request {i} differs from the oracle's image"
