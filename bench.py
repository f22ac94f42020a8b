"""Benchmark: Mpaths/s of the wavefront path tracer at 1080p on one MI355X.

Workload (one step = one full render): the material scene — ground plane,
25 subdivision-4 icospheres (128,010 triangles) cycling over textured
Lambertian, glass, GGX reflection, GGX transmission and normal-mapped
Lambertian, 4 emissive quads, environment 0.1 — at 1920x1080, 256 spp,
maximum depth 16, NEE + MIS + Russian roulette from depth 3, every path run to
completion (no 8-iteration cap), FP32. Scene and BVH are built before the
timed region; inputs are resident in HBM.

Prints one JSON line (bench contract) with a roofline object for the
dominant kernel (closest-hit traversal: algorithmic bytes per launch over
the HIP-event launch time, and the HBM traffic measured with rocprofv3
counters), per-kernel times and item counts, a `large` object with the same
figures for the incoherent 10.24M-triangle scene (2,000 displaced
icospheres, 3840x2160, 16 spp, depth 8; --no-large skips it), and a
CPU-baseline object (the oracle, OpenMP on the host cores granted to the
job, on a sample of whole 1-spp frames of the same workload, with the host's
nproc and CPU model).

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--spp 256]

--workload large measures the 10.24M-triangle scene as the main line; the
default workload is the one the headline metric is quoted on.

--workload serving measures the continuous-batching server: the Cornell,
material and large scenes resident on the device, a seeded stream of
requests (random camera; 256x256 / 512x512 / 1024x768 / 1920x1080;
1 / 4 / 16 / 64 spp; depth 8). It reports aggregate Mpaths/s with every
request submitted at once against the same requests rendered one after
another, and p50/p99 latency (completion - arrival) at a fixed Poisson
arrival rate for the server and for serial FIFO rendering, and checks that
each served image equals its stand-alone render bit for bit.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
TRI_BYTES = 48     # leaf-ordered triangle record (3 x float4)
# closest-hit per ray: queue index 4 + ray (origin,tMin | dir,tMax) 32 + hit
# record 8 + shade category 1. The resolve epilogue's path-state accesses on a
# miss (throughput, radiance and state read, radiance and result written: 80 B)
# are not counted, so the figure is a lower bound.
CLOSEST_FIXED = 4 + 32 + 8 + 1
# any-hit per ray: queue index 4 + ray 32 + category 1 (an unoccluded ray's
# radiance and contribution, 48 B, not counted)
ANY_FIXED = 4 + 32 + 1


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["material", "large", "serving", "denoise", "denoise-infer"],
                    default="material")
    ap.add_argument("--spp", type=int, default=None)
    ap.add_argument("--width", type=int, default=None)
    ap.add_argument("--height", type=int, default=None)
    ap.add_argument("--max-paths", type=int, default=None,
                    help="paths in flight (default 256M = ~46 GB of HBM; 128M for serving)")
    ap.add_argument("--max-chunk", type=int, default=None, help="paths per admitted chunk (default: whole pool; "
                    "8M for serving)")
    ap.add_argument("--requests", type=int, default=256, help="serving: number of requests")
    ap.add_argument("--rate", type=float, default=40.0, help="serving: Poisson arrival rate, requests/s")
    ap.add_argument("--rate-sweep", type=str, default="",
                    help="serving: further arrival rates (comma-separated req/s), each an open-loop run of its own "
                    "reported with serial FIFO latency under the same arrivals")
    ap.add_argument("--round-robin", action="store_true",
                    help="serving: one chunk per pending request per round (default: in-order admission)")
    ap.add_argument("--express-max-paths", type=int, default=1 << 22,
                    help="serving: requests of at most this many paths take the express lane (0 = off)")
    ap.add_argument("--fused-max-paths", type=int, default=1 << 20,
                    help="serving: requests of at most this many paths render in one fused launch per chunk")
    ap.add_argument("--serial-fused-max-paths", type=int, default=0,
                    help="serving: the same threshold for the serial per-request renders")
    ap.add_argument("--serial-max-paths", type=int, default=1 << 27,
                    help="serving: paths in flight of the serial per-request renders")
    ap.add_argument("--express-pool", type=int, default=1 << 23, help="serving: express lane paths in flight")
    ap.add_argument("--backfill-paths", type=int, default=0,
                    help="serving: pool entries the requests behind the oldest unfinished one may hold while it has "
                    "paths in flight (0 = no limit)")
    ap.add_argument("--express-iters", type=int, default=32,
                    help="serving: express lane iterations per server step")
    ap.add_argument("--express-overlap", action="store_true",
                    help="serving: express iterations run while the main iteration is on the device "
                    "(default: before it)")
    ap.add_argument("--policy", choices=["fifo", "shortest"], default="fifo",
                    help="serving: admission order of pending requests")
    ap.add_argument("--dn-scenes", type=int, default=24, help="denoise: training scenes")
    ap.add_argument("--dn-cams", type=int, default=48, help="denoise: cameras per training scene")
    ap.add_argument("--dn-target-spp", type=int, default=2048, help="denoise: spp of the training targets")
    ap.add_argument("--dn-steps", type=int, default=60000, help="denoise: training steps (16 crops of 128^2)")
    ap.add_argument("--dn-loss", choices=["relmse_split", "relmse", "relmse_pred", "log"], default="relmse_split",
                    help="denoise: training loss")
    ap.add_argument("--dn-lr", type=float, default=1e-3, help="denoise: peak learning rate")
    ap.add_argument("--dn-footprint", type=str, default=None,
                    help="denoise: train on 256^2 windows at the pixel scale of a WxH frame (e.g. 1920x1080); "
                    "default: whole 256^2 frames")
    ap.add_argument("--dn-load", type=str, default=None,
                    help="denoise: start from these weights (with --dn-steps 0: evaluate them only)")
    ap.add_argument("--dn-input-spp", type=str, default="4,16,64,256",
                    help="denoise: sample counts of the training inputs (comma-separated)")
    ap.add_argument("--dn-head", choices=["residual", "kernel"], default="residual",
                    help="denoise: network output, a log-colour correction or a 7x7 filter per pixel")
    ap.add_argument("--dn-no-noise", action="store_true",
                    help="denoise: train without the per-pixel noise input (channel 10 left at 0)")
    ap.add_argument("--dn-eval-size", type=str, default="1920x1080", help="denoise: evaluation resolution")
    ap.add_argument("--dn-ref-spp", type=int, default=4096, help="denoise: spp of the evaluation reference")
    ap.add_argument("--dn-save", type=str, default=None, help="denoise: write the trained weights here")
    ap.add_argument("--dn-checkpoint", type=str, default=None,
                    help="denoise: training checkpoint (.npz: weights, Adam state, crop generator, next step)")
    ap.add_argument("--dn-checkpoint-every", type=int, default=5000, help="denoise: steps between checkpoints")
    ap.add_argument("--dn-resume", type=str, default=None, help="denoise: continue training from this checkpoint")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="target wall time of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-large", action="store_true",
                    help="material workload: skip the large-scene sub-measurement")
    ap.add_argument("--large-steps", type=int, default=2, help="timed renders of the large-scene sub-measurement")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="after the bench, run a plain device copy of this many MiB (calibrates HBM counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (one-GPU build)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import offlinerenderer_amd as wr
    from offlinerenderer_amd import scenes

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    wr.Context(wr.ContextConfig(deviceId=0, logLevel=0))
    if args.workload == "serving":
        serving(args)
        return
    if args.workload == "denoise":
        denoise(args)
        return
    if args.workload == "denoise-infer":
        denoise_infer(args)
        return
    if args.max_paths is None:
        args.max_paths = 1 << 28
    if args.workload == "material":
        W, H, spp = args.width or 1920, args.height or 1080, args.spp or 256
    else:
        W, H, spp = args.width or 3840, args.height or 2160, args.spp or 16
    m = measure_render(args, args.workload, W, H, spp, args.steps, args.warmup)
    line = m["line"]
    if args.workload == "material" and not args.no_large:
        # the 10.24M-triangle scene (the traversal target is defined on it)
        lg = measure_render(args, "large", 3840, 2160, 16, args.large_steps, 1)
        line["large"] = {k: lg["line"][k] for k in ("value", "unit", "ms_per_step", "steps", "config",
                                                     "mrays_per_s", "bvh_build_ms", "kernel_ms_per_step",
                                                     "kernel_items_per_step", "roofline")}
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(m["desc"], W, H, args.cpu_seconds)
    line["cpu_baseline"] = cpu
    print(json.dumps(line), flush=True)


def measure_render(args, workload_kind, W, H, spp, steps, warmup):
    """Timed renders of one workload: Mpaths/s, per-kernel times and items,
    and the closest-hit roofline (algorithmic and measured traffic)."""
    import numpy as np
    import torch

    import offlinerenderer_amd as wr
    from offlinerenderer_amd import scenes

    if workload_kind == "material":
        desc = scenes.material_scene(W, H, spp)
        workload = f"material_scene_{W}x{H}_{spp}spp_depth{desc.maxBounces}"
    else:
        desc = scenes.large_scene(W, H, spp)
        workload = f"large_scene_{desc.numTris}tris_{W}x{H}_{spp}spp_depth{desc.maxBounces}"
    log(f"[bench] {workload}")
    scene = wr.Scene()
    desc.build(scene)
    scene.finalize()
    bvh = scene.getBVH()
    log(f"[bench] {desc.numTris} tris, BVH {bvh.numNodes} nodes depth {bvh.maxDepth}, "
        f"build {scene.getBuildSeconds() * 1e3:.1f} ms")
    bvh_build_ms = scene.getBuildSeconds() * 1e3
    cam = wr.Camera(**desc.camera)
    prm = wr.RenderParams(width=W, height=H, spp=spp, maxBounces=desc.maxBounces,
                          russianRouletteDepth=desc.rrDepth)
    r = wr.Renderer()

    # fetch counts per ray (counting kernels, 1 spp, untimed)
    r.setWavefrontConfig(wr.WavefrontConfig(iterationCap=0, maxPathsInFlight=args.max_paths, countTraversal=True))
    cprm = wr.RenderParams(width=W, height=H, spp=1, maxBounces=desc.maxBounces, russianRouletteDepth=desc.rrDepth)
    r.render(scene, cam, cprm)
    cs = r.getStats()
    nodes_per_ray = cs["closestNodeFetches"] / max(cs["closestRays"], 1)
    tris_per_ray = cs["closestTriFetches"] / max(cs["closestRays"], 1)
    any_nodes = cs["anyNodeFetches"] / max(cs["shadowRays"], 1)
    any_tris = cs["anyTriFetches"] / max(cs["shadowRays"], 1)
    spills_per_ray = cs["closestStackSpills"] / max(cs["closestRays"], 1)
    any_spills = cs["anyStackSpills"] / max(cs["shadowRays"], 1)
    log(f"[bench] closest: {nodes_per_ray:.2f} nodes, {tris_per_ray:.2f} tris, {spills_per_ray:.4f} stack spills "
        f"per ray; any: {any_nodes:.2f} nodes, {any_tris:.2f} tris, {any_spills:.4f} spills per ray")

    r.setWavefrontConfig(wr.WavefrontConfig(iterationCap=0, maxPathsInFlight=args.max_paths,
                                            maxChunkPaths=args.max_chunk or 0, timeKernels=True))
    for i in range(warmup):
        r.render(scene, cam, prm)
        log(f"[bench] warmup {i + 1}/{warmup}: {r.getStats()['seconds']:.3f} s")
    totals = {k: 0.0 for k in ("msTraceClosest", "msTraceAny", "msShade", "msClassify", "msPartition",
                                "msGenerate", "msAccumulate")}
    items = {k: 0 for k in ("closestRays", "shadowRays", "classifyItems", "partitionItems", "shadeItems")}
    launches_c = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        img = r.render(scene, cam, prm)
        st = r.getStats()
        for k in totals:
            totals[k] += st[k]
        for k in items:
            items[k] += st[k]
        launches_c += st["launchesTraceClosest"]
        log(f"[bench] step {i + 1}/{steps}: {st['seconds']:.3f} s, "
            f"{st['paths'] / st['seconds'] / 1e6:.1f} Mpaths/s")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    assert np.isfinite(img).all()
    rays_c, rays_s = items["closestRays"], items["shadowRays"]
    paths = W * H * spp * steps
    mpaths = paths / elapsed / 1e6
    mrays = (rays_c + rays_s) / elapsed / 1e6

    bytes_per_ray = CLOSEST_FIXED + nodes_per_ray * bvh.nodeBytes + tris_per_ray * TRI_BYTES
    closest_s = totals["msTraceClosest"] / 1e3
    achieved = bytes_per_ray * rays_c / closest_s / 1e9
    avg_launch_ms = totals["msTraceClosest"] / max(launches_c, 1)
    any_bytes_per_ray = ANY_FIXED + any_nodes * bvh.nodeBytes + any_tris * TRI_BYTES
    any_gbs = any_bytes_per_ray * rays_s / max(totals["msTraceAny"] / 1e3, 1e-9) / 1e9

    if args.copy_probe_mib and workload_kind == args.workload:
        nbytes = args.copy_probe_mib << 20
        src = torch.empty(nbytes, dtype=torch.uint8, device="cuda").fill_(1)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] copy probe: 3 x {nbytes} bytes read + written")

    # HBM bytes per launch from the rocprofv3 FETCH_SIZE/WRITE_SIZE runs
    # (profiles/hbm_traffic_pmc.md): measured bytes per ray x rays per launch
    traffic, traffic_src = None, None
    tfile = ROOT / "profiles" / "hbm_traffic.json"
    if tfile.exists():
        tj = json.loads(tfile.read_text())
        t = tj.get(workload_kind)
        if t:
            traffic = round(t["hbm_bytes_per_ray"] * rays_c / max(launches_c, 1))
            traffic_src = f"profiles/hbm_traffic.json ({tj.get('measured_at', 'unrecorded build')})"
    avg_launch_s = avg_launch_ms / 1e3
    measured_gbs = traffic / avg_launch_s / 1e9 if traffic and avg_launch_s > 0 else None

    line = {
        "metric": "Mpaths/s @1080p; time-to-relMSE<=1e-2 learned vs plain; p50/p99 req latency",
        "value": round(mpaths, 3),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(elapsed / steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": f"synthetic (seeded {workload_kind} scene generator)",
        "config": {
            "workload": workload,
            "width": W, "height": H, "spp": spp, "max_depth": desc.maxBounces, "triangles": desc.numTris,
            "iteration_cap": "none (paths run to completion)", "paths_in_flight": args.max_paths,
            "chunk_paths": args.max_chunk or args.max_paths,
        },
        "mrays_per_s": round(mrays, 3),
        "bvh_build_ms": round(bvh_build_ms, 2),
        "kernel_ms_per_step": {k: round(v / steps, 3) for k, v in totals.items()},
        "kernel_items_per_step": {k: round(v / steps) for k, v in items.items()},
        "roofline": {
            "kernel": "k_trace<closest>",
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "frac_basis": "algorithmic bytes (every node and triangle fetch, pre-cache)",
            "traffic": traffic,
            "traffic_unit": "bytes per launch (HBM, rocprofv3 FETCH_SIZE+WRITE_SIZE, calibrated)",
            "traffic_source": traffic_src,
            "measured_gbs": round(measured_gbs, 2) if measured_gbs else None,
            "measured_frac": round(measured_gbs / HBM_PEAK_GBS, 4) if measured_gbs else None,
            "limiter": "node-test instruction issue and the latency of dependent node fetches, with under half "
                       "of each wave's lanes active, not HBM bandwidth: the "
                       "measured HBM traffic is below the algorithmic bytes (cache hits) and below 8 TB/s",
            "algorithmic_bytes_per_launch": round(bytes_per_ray * rays_c / max(launches_c, 1)),
            "bytes_per_ray": round(bytes_per_ray, 2),
            "nodes_per_ray": round(nodes_per_ray, 3),
            "tris_per_ray": round(tris_per_ray, 3),
            "stack_spills_per_ray": round(spills_per_ray, 5),
            "any_stack_spills_per_ray": round(any_spills, 5),
            "avg_launch_ms": round(avg_launch_ms, 4),
            "rays_per_launch": round(rays_c / max(launches_c, 1), 1),
            "any_hit_gbs": round(any_gbs, 2),
        },
    }
    return {"line": line, "desc": desc}


def serving(args):
    """Continuous batching against serial per-request rendering."""
    import hashlib

    import numpy as np

    import offlinerenderer_amd as wr
    from offlinerenderer_amd import scenes

    pool = args.max_paths or (1 << 27)
    chunk = args.max_chunk or (1 << 23)
    descs = {"cornell": scenes.cornell(), "material": scenes.material_scene(), "large": scenes.large_scene()}
    resident = {}
    for name, d in descs.items():
        s = wr.Scene()
        d.build(s)
        s.finalize()
        resident[name] = s
        log(f"[serve] scene {name}: {d.numTris} tris, BVH build {s.getBuildSeconds() * 1e3:.1f} ms")
    reqs = scenes.serving_requests(args.requests, seed=7)
    jobs = [(resident[q["scene"]], wr.Camera(**q["camera"]),
             wr.RenderParams(width=q["width"], height=q["height"], spp=q["spp"], maxBounces=q["maxBounces"],
                             russianRouletteDepth=q["rrDepth"])) for q in reqs]
    paths = np.array([p.width * p.height * p.spp for _, _, p in jobs], dtype=np.float64)
    total_paths = float(paths.sum())
    def arrivals_at(rate):
        # the same unit-rate exponential draws at every rate: only the time scale changes
        a = np.cumsum(np.random.default_rng(11).exponential(1.0, size=len(jobs))) / rate
        return a - a[0]

    arrivals = arrivals_at(args.rate)

    def digest(img):
        return hashlib.sha1(np.ascontiguousarray(img).tobytes()).hexdigest()

    # serial: one request after another, each on a pool that holds it whole
    # (128M paths, the largest request is 1920x1080 x 64 spp = 133M)
    serial_pool = args.serial_max_paths
    r = wr.Renderer()
    r.setWavefrontConfig(wr.WavefrontConfig(iterationCap=0, maxPathsInFlight=serial_pool,
                                            fusedMaxPaths=args.serial_fused_max_paths))
    r.render(*jobs[0])  # warmup
    service, ref = np.zeros(len(jobs)), []
    for i, job in enumerate(jobs):
        t0 = time.perf_counter()
        img = r.render(*job)
        service[i] = time.perf_counter() - t0
        ref.append(digest(img))
    del r
    serial_mpaths = total_paths / service.sum() / 1e6
    log(f"[serve] serial: {service.sum():.2f} s busy, {serial_mpaths:.1f} Mpaths/s")
    def fifo_latency(arrive):
        # serial FIFO latency under these arrivals (one request at a time)
        finish, t = np.zeros(len(jobs)), 0.0
        for i in range(len(jobs)):
            t = max(t, arrive[i]) + service[i]
            finish[i] = t
        return finish - arrive

    serial_lat = fifo_latency(arrivals)

    def run_server(arrive):
        srv = wr.RenderServer(wr.ServerConfig(maxPathsInFlight=pool, maxChunkPaths=chunk,
                                              shortestFirst=args.policy == "shortest",
                                              roundRobin=args.round_robin,
                                              expressMaxPaths=args.express_max_paths,
                                              expressPathsInFlight=args.express_pool,
                                              expressIterations=args.express_iters,
                                              expressOverlap=args.express_overlap,
                                              fusedMaxPaths=args.fused_max_paths,
                                              backfillPaths=args.backfill_paths))
        ids, done_t, images = [None] * len(jobs), np.zeros(len(jobs)), [None] * len(jobs)
        open_ids = {}
        t0 = srv.clock()
        nxt, steps = 0, 0
        while nxt < len(jobs) or open_ids:
            now = srv.clock() - t0
            while nxt < len(jobs) and arrive[nxt] <= now:
                ids[nxt] = srv.submit(*jobs[nxt])
                open_ids[ids[nxt]] = nxt
                nxt += 1
            if not open_ids:
                time.sleep(max(0.0, arrive[nxt] - (srv.clock() - t0)))
                continue
            srv.step()
            steps += 1
            for rid in srv.takeCompleted():
                i = open_ids.pop(rid)
                done_t[i] = srv.status(rid)["doneTime"] - t0
                images[i] = srv.fetch(rid)
        stats = srv.getStats()
        # hashing stays outside the served loop: the GPU would idle meanwhile
        equal = sum(digest(img) == ref[i] for i, img in enumerate(images))
        return done_t, equal, steps, stats

    # saturation: every request at t = 0
    done_sat, eq_sat, steps_sat, st_sat = run_server(np.zeros(len(jobs)))
    makespan = float(done_sat.max())
    server_mpaths = total_paths / makespan / 1e6
    log(f"[serve] server, all at once: {makespan:.2f} s, {server_mpaths:.1f} Mpaths/s, {steps_sat} steps")
    # open loop at the fixed arrival rate
    done_ol, eq_ol, steps_ol, _ = run_server(arrivals)
    server_lat = done_ol - arrivals
    log(f"[serve] server at {args.rate} req/s: p50 {np.percentile(server_lat, 50) * 1e3:.0f} ms, "
        f"p99 {np.percentile(server_lat, 99) * 1e3:.0f} ms")

    def pct(a):
        return {"p50_ms": round(float(np.percentile(a, 50)) * 1e3, 1),
                "p99_ms": round(float(np.percentile(a, 99)) * 1e3, 1),
                "mean_ms": round(float(a.mean()) * 1e3, 1)}

    sweep = []
    for rate in [float(v) for v in args.rate_sweep.split(",") if v.strip()]:
        arr = arrivals_at(rate)
        done_r, eq_r, _, _ = run_server(arr)
        lat_r, fifo_r = done_r - arr, fifo_latency(arr)
        sweep.append({"rate_per_s": rate,
                      "offered_load": round(float(paths.sum() / arr[-1]) / (serial_mpaths * 1e6), 3),
                      "server": pct(lat_r), "serial_fifo": pct(fifo_r),
                      "mixed_equals_alone": f"{eq_r}/{len(jobs)}"})
        log(f"[serve] server at {rate:g} req/s: p50/p99 {sweep[-1]['server']['p50_ms']:.0f}/"
            f"{sweep[-1]['server']['p99_ms']:.0f} ms, serial FIFO {sweep[-1]['serial_fifo']['p50_ms']:.0f}/"
            f"{sweep[-1]['serial_fifo']['p99_ms']:.0f} ms")
        assert eq_r == len(jobs), "a served image differs from its stand-alone render"

    line = {
        "metric": "Mpaths/s @1080p; time-to-relMSE<=1e-2 learned vs plain; p50/p99 req latency",
        "value": round(server_mpaths, 3),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": round(makespan * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded request stream over the Cornell, material and large scenes)",
        "config": {
            "workload": f"serving_{len(jobs)}req_poisson{args.rate:g}",
            "requests": len(jobs), "total_paths": int(total_paths), "resolutions": list(scenes.SERVING_RESOLUTIONS),
            "spp": list(scenes.SERVING_SPP), "max_depth": 8, "arrival_rate_per_s": args.rate,
            "paths_in_flight": pool, "chunk_paths": chunk, "serial_paths_in_flight": serial_pool,
            "fused_max_paths": args.fused_max_paths, "serial_fused_max_paths": args.serial_fused_max_paths,
            "admission": args.policy + (" round-robin" if args.round_robin else " in order"),
            "backfill_paths": args.backfill_paths or None,
            "express_lane": ({"max_paths": args.express_max_paths, "paths_in_flight": args.express_pool,
                              "iterations_per_step": args.express_iters,
                              "schedule": "overlapped" if args.express_overlap else "time-sliced"}
                             if args.express_max_paths else None),
        },
        "serial_mpaths_per_s": round(serial_mpaths, 3),
        "speedup_vs_serial": round(server_mpaths / serial_mpaths, 3),
        "latency_server": pct(server_lat),
        "latency_serial_fifo": pct(serial_lat),
        # the same percentiles over requests of at most 4M paths (10 of the
        # 16 size classes, ~8% of the paths) and over the larger ones
        "latency_by_size": {
            cls: {"requests": int(m.sum()), "server": pct(server_lat[m]), "serial_fifo": pct(serial_lat[m])}
            for cls, m in (("le_4M_paths", paths <= (1 << 22)), ("gt_4M_paths", paths > (1 << 22)))
            if m.any()},
        "offered_load": round(float(paths.sum() / arrivals[-1]) / (serial_mpaths * 1e6), 3) if arrivals[-1] > 0 else None,
        "rate_sweep": sweep or None,
        "server_steps": {"saturated": steps_sat, "open_loop": steps_ol},
        "fused_paths_saturated": int(st_sat["fusedPaths"]),
        "mixed_equals_alone": {"saturated": f"{eq_sat}/{len(jobs)}", "open_loop": f"{eq_ol}/{len(jobs)}"},
        "mrays_per_s_saturated": round((st_sat["closestRays"] + st_sat["shadowRays"]) / makespan / 1e6, 3),
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    assert eq_sat == len(jobs) and eq_ol == len(jobs), "a served image differs from its stand-alone render"


def denoise(args):
    """Learned denoiser: train on the tracer's own data, then time to
    relMSE <= 1e-2 at the evaluation resolution, learned vs plain."""
    import numpy as np

    from offlinerenderer_amd import denoise_train as dt
    from offlinerenderer_amd import denoiser

    footprint = tuple(int(v) for v in args.dn_footprint.split("x")) if args.dn_footprint else None
    net = denoiser.UNet(seed=1, head=args.dn_head)
    if args.dn_load:
        net.load(args.dn_load)
    t_data = t_train = 0.0
    losses = [float("nan")]
    if args.dn_steps > 0:
        t0 = time.perf_counter()
        data = dt.build_dataset(num_scenes=args.dn_scenes, cams_per_scene=args.dn_cams,
                                noisy_spps=tuple(int(v) for v in args.dn_input_spp.split(",")),
                                spp_target=args.dn_target_spp, footprint=footprint, log=log)
        t_data = time.perf_counter() - t0
        if args.dn_no_noise:
            data["noise"] = None  # ablation: input channel 10 stays 0
        t0 = time.perf_counter()
        losses = dt.train(net, data, steps=args.dn_steps, lr=args.dn_lr, loss=args.dn_loss, log=log,
                          checkpoint=args.dn_checkpoint, checkpoint_every=args.dn_checkpoint_every,
                          resume=args.dn_resume)
        t_train = time.perf_counter() - t0
        del data
    if args.dn_save:
        net.save(args.dn_save)
    W, H = (int(v) for v in args.dn_eval_size.split("x"))
    res = dt.evaluate(net, width=W, height=H, ref_spp=args.dn_ref_spp, log=log)
    if res["learned_seconds_to_bar"] and res["plain_seconds_to_bar"]:
        speedup = res["plain_seconds_to_bar"] / res["learned_seconds_to_bar"]
        basis = f"time to relMSE<=1e-2 (learned at {res['learned_spp_to_bar']} spp + network)"
    elif "plain_seconds_equal_quality" in res:
        speedup, basis = res["plain_seconds_equal_quality"] / res["learned_seconds"], "time to equal relMSE at 4 spp"
    else:
        speedup, basis = None, "not comparable"
    peak = 2500.0  # MI355X dense bf16 MFMA, TFLOP/s
    line = {
        "metric": "Mpaths/s @1080p; time-to-relMSE<=1e-2 learned vs plain; p50/p99 req latency",
        "value": None if speedup is None else round(speedup, 3),
        "unit": "x (plain time / learned time)",
        "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": round(res["learned_seconds"] * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (network), f32 (tracer)",
        "data": "synthetic (tracer-generated: seeded large-scene generator, random cameras)",
        "config": {
            "workload": f"denoise_{W}x{H}_4spp_unet",
            "train_views": args.dn_scenes * args.dn_cams, "train_input_spp": [int(v) for v in args.dn_input_spp.split(",")], "train_res": 256,
            "train_target_spp": args.dn_target_spp,
            "train_steps": args.dn_steps, "batch": "16 x 128^2 crops", "params": net.num_params,
            "train_view_kind": ("256^2 windows at the pixel scale of " + args.dn_footprint) if footprint
            else "whole 256^2 frames",
            "weights_loaded": args.dn_load,
            "loss": args.dn_loss,
            "head": args.dn_head,
            "noise_input": not args.dn_no_noise,
            "eval_reference_spp": args.dn_ref_spp,
        },
        "speedup_basis": basis,
        "learned": {"rel_mse": round(res["learned_rel_mse"], 5), "ms": round(res["learned_seconds"] * 1e3, 2)},
        "learned_sweep": [{k: (round(v, 5) if isinstance(v, float) else v) for k, v in p.items()}
                          for p in res["learned_sweep"]],
        "learned_spp_to_bar": res["learned_spp_to_bar"],
        "equal_quality_speedup_at_4spp": (round(res["plain_seconds_equal_quality"] / res["learned_seconds"], 2)
                                          if "plain_seconds_equal_quality" in res else None),
        "noisy_4spp": {"rel_mse": round(res["noisy_rel_mse"], 5), "ms": round(res["noisy_seconds"] * 1e3, 2)},
        "plain": [{k: (round(v, 5) if isinstance(v, float) else v) for k, v in p.items()} for p in res["plain"]],
        "plain_spp_to_bar": res["plain_spp_to_bar"],
        "plain_spp_equal_quality": res.get("plain_spp_equal_quality"),
        "train": {"data_s": round(t_data, 1), "train_s": round(t_train, 1),
                  "loss_first100": round(float(np.mean(losses[:100])), 5),
                  "loss_last100": round(float(np.mean(losses[-100:])), 5)},
        "roofline": {"kernel": "U-Net forward (15 bf16 MFMA convolutions)", "bound": "mfma",
                     "achieved": round(res["inference_tflops"], 2), "peak": peak, "unit": "TFLOP/s",
                     "frac": round(res["inference_tflops"] / peak, 4), "traffic": None,
                     "inference_ms": round(res["inference_ms"], 3)},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def denoise_infer(args):
    """The learned path alone with the shipped weights: 4-spp render with
    guides + U-Net inference on a held-out large scene, timed per step."""
    import numpy as np
    import torch

    import offlinerenderer_amd as wr
    from offlinerenderer_amd import denoise_train as dt
    from offlinerenderer_amd import denoiser

    W, H = (int(v) for v in args.dn_eval_size.split("x"))
    d, s = dt.build_scene(9001, W, H)
    cam = wr.Camera(**dt.large_camera(np.random.default_rng(9001), W, H))
    r = wr.Renderer()
    r.setWavefrontConfig(wr.WavefrontConfig(iterationCap=0, maxPathsInFlight=1 << 28))
    p = wr.RenderParams(width=W, height=H, spp=4, maxBounces=d.maxBounces, russianRouletteDepth=d.rrDepth)
    p.denoiser.enabled = True
    for _ in range(args.warmup):
        r.render(s, cam, p)
    times = []
    for i in range(args.steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r.render(s, cam, p)
        times.append(time.perf_counter() - t0)
        log(f"[denoise-infer] step {i + 1}/{args.steps}: {times[-1] * 1e3:.1f} ms")
    # the network alone on device-resident layers
    net = denoiser.UNet()
    net.load(str(wr.LIB_PATH.parent.parent / "models" / "denoiser_unet.f32"))
    q = wr.RenderParams(width=W, height=H, spp=4, maxBounces=d.maxBounces, russianRouletteDepth=d.rrDepth)
    dev = [torch.from_numpy(a).cuda().unsqueeze(0) for a in r.renderWithGuides(s, cam, q, noise=True)]
    noise = dev.pop()
    out = torch.empty_like(dev[0])
    net.set_sample_counts([4])
    net.set_noise(noise)
    net.forward(*dev, out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        net.forward(*dev, out)
    torch.cuda.synchronize()
    t_inf = (time.perf_counter() - t0) / args.steps
    tflops = denoiser.flops_per_pixel() * W * H / t_inf / 1e12
    print(json.dumps({
        "metric": "Mpaths/s @1080p; time-to-relMSE<=1e-2 learned vs plain; p50/p99 req latency",
        "value": round(float(np.median(times)) * 1e3, 3), "unit": "ms per denoised 4-spp frame",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(float(np.median(times)) * 1e3, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (network), f32 (tracer)",
        "data": "synthetic (seeded large-scene generator, shipped weights)",
        "config": {"workload": f"denoise_infer_{W}x{H}_4spp"},
        "roofline": {"kernel": "U-Net forward (15 bf16 MFMA convolutions)", "bound": "mfma",
                     "achieved": round(tflops, 2), "peak": 2500.0, "unit": "TFLOP/s", "frac": round(tflops / 2500.0, 4),
                     "traffic": None, "inference_ms": round(t_inf * 1e3, 3)},
        "cpu_baseline": None,
    }), flush=True)


def host_cpu():
    """nproc, the cores this process may run on, and the CPU model (lscpu)."""
    import subprocess

    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in out.splitlines():
            if ln.startswith("Model name:"):
                model = ln.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    if model is None and Path("/proc/cpuinfo").exists():
        for ln in Path("/proc/cpuinfo").read_text().splitlines():
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    return {"nproc": os.cpu_count(), "affinity_cores": len(os.sched_getaffinity(0)), "model": model,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(desc, W, H, target_s):
    """Oracle (the CPU restatement of libWR's kernel math) on the host cores
    granted to this job: OMP_NUM_THREADS when the host sets it (the GPU hosts
    give one GPU's job 16 threads), else every core of the affinity mask."""
    from oracle import pyoracle as po

    hc = host_cpu()
    threads = hc["affinity_cores"]
    if hc["omp_num_threads"] and hc["omp_num_threads"].isdigit():
        threads = max(1, min(threads, int(hc["omp_num_threads"])))
    log(f"[bench] CPU baseline: oracle with host BVH, {threads} threads on {hc['model']} "
        f"(nproc {hc['nproc']}, affinity {hc['affinity_cores']})")
    os_ = po.build_oracle_scene(desc, bvh=True)
    prm = po.params_in(W, H, desc.maxBounces, desc.rrDepth, iterationCap=0)
    done_spp, secs = 0, 0.0
    while done_spp < 32 and (secs < target_s or done_spp == 0):
        t0 = time.perf_counter()
        po.render_image(os_, desc.camera, prm, 1, s0=done_spp, use_bvh=True, threads=threads)
        secs += time.perf_counter() - t0
        done_spp += 1
        log(f"[bench] CPU baseline: {done_spp} spp in {secs:.2f} s")
    paths = W * H * done_spp
    return {"value": round(paths / secs / 1e6, 4), "unit": "Mpaths/s", "cores": threads, "kind": "port",
            "sample": f"{done_spp} spp of the same {W}x{H} frame (samples 0..{done_spp - 1}), full depth "
                      f"{desc.maxBounces}, {secs:.1f} s",
            "host": hc}


if __name__ == "__main__":
    main()
