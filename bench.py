"""Benchmark: Mpaths/s of the wavefront path tracer at 1080p on one MI355X.

Workload (one step = one full render): the material scene — ground plane,
25 subdivision-4 icospheres (128,010 triangles) cycling over textured
Lambertian, glass, GGX reflection, GGX transmission and normal-mapped
Lambertian, 4 emissive quads, environment 0.1 — at 1920x1080, 256 spp,
maximum depth 16, NEE + MIS + Russian roulette from depth 3, every path run to
completion (no 8-iteration cap), FP32. Scene and BVH are built before the
timed region; inputs are resident in HBM.

Prints one JSON line (bench contract) with a roofline object for the
dominant kernel (closest-hit traversal) and a CPU-baseline object (the
oracle, OpenMP on the host cores, on a 1-spp sample of the same frame).

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--spp 256]

--workload large measures the incoherent 10.24M-triangle scene instead
(2,000 displaced icospheres, 3840x2160, 16 spp, depth 8); the default
workload is the one the headline metric is quoted on.
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
NODE_BYTES = 128   # 4-wide node record
TRI_BYTES = 48     # leaf-ordered triangle record (3 x float4)
# closest-hit per ray: queue index 4 + ray (origin,tMin | dir,tMax) 32 + hit record 8
CLOSEST_FIXED = 4 + 32 + 8
# any-hit per ray: queue index 4 + ray 32 + occlusion byte 1
ANY_FIXED = 4 + 32 + 1


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["material", "large"], default="material")
    ap.add_argument("--spp", type=int, default=None)
    ap.add_argument("--width", type=int, default=None)
    ap.add_argument("--height", type=int, default=None)
    ap.add_argument("--max-paths", type=int, default=1 << 28, help="paths in flight (256M = ~46 GB of HBM)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="target wall time of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
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
    if args.workload == "material":
        W, H, spp = args.width or 1920, args.height or 1080, args.spp or 256
        desc = scenes.material_scene(W, H, spp)
        workload = f"material_scene_{W}x{H}_{spp}spp_depth{desc.maxBounces}"
    else:
        W, H, spp = args.width or 3840, args.height or 2160, args.spp or 16
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
    log(f"[bench] closest: {nodes_per_ray:.2f} nodes, {tris_per_ray:.2f} tris per ray; "
        f"any: {any_nodes:.2f} nodes, {any_tris:.2f} tris per ray")

    r.setWavefrontConfig(wr.WavefrontConfig(iterationCap=0, maxPathsInFlight=args.max_paths, timeKernels=True))
    for i in range(args.warmup):
        r.render(scene, cam, prm)
        log(f"[bench] warmup {i + 1}/{args.warmup}: {r.getStats()['seconds']:.3f} s")
    totals = {k: 0.0 for k in ("msTraceClosest", "msTraceAny", "msShade", "msClassify", "msPartition",
                                "msGenerate", "msAccumulate")}
    rays_c = rays_s = launches_c = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        img = r.render(scene, cam, prm)
        st = r.getStats()
        for k in totals:
            totals[k] += st[k]
        rays_c += st["closestRays"]
        rays_s += st["shadowRays"]
        launches_c += st["launchesTraceClosest"]
        log(f"[bench] step {i + 1}/{args.steps}: {st['seconds']:.3f} s, "
            f"{st['paths'] / st['seconds'] / 1e6:.1f} Mpaths/s")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    assert np.isfinite(img).all()
    paths = W * H * spp * args.steps
    mpaths = paths / elapsed / 1e6
    mrays = (rays_c + rays_s) / elapsed / 1e6

    bytes_per_ray = CLOSEST_FIXED + nodes_per_ray * NODE_BYTES + tris_per_ray * TRI_BYTES
    closest_s = totals["msTraceClosest"] / 1e3
    achieved = bytes_per_ray * rays_c / closest_s / 1e9
    avg_launch_ms = totals["msTraceClosest"] / max(launches_c, 1)
    any_bytes_per_ray = ANY_FIXED + any_nodes * NODE_BYTES + any_tris * TRI_BYTES
    any_gbs = any_bytes_per_ray * rays_s / max(totals["msTraceAny"] / 1e3, 1e-9) / 1e9

    if args.copy_probe_mib:
        nbytes = args.copy_probe_mib << 20
        src = torch.empty(nbytes, dtype=torch.uint8, device="cuda").fill_(1)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] copy probe: 3 x {nbytes} bytes read + written")

    # HBM bytes per launch from the rocprofv3 FETCH_SIZE/WRITE_SIZE runs
    # (profiles/hbm_traffic_pmc.md): measured bytes per ray x rays per launch
    traffic = None
    tfile = ROOT / "profiles" / "hbm_traffic.json"
    if tfile.exists():
        t = json.loads(tfile.read_text()).get(args.workload)
        if t:
            traffic = round(t["hbm_bytes_per_ray"] * rays_c / max(launches_c, 1))

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(desc, W, H, args.cpu_seconds)

    line = {
        "metric": "Mpaths/s @1080p; time-to-relMSE<=1e-2 learned vs plain; p50/p99 req latency",
        "value": round(mpaths, 3),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": f"synthetic (seeded {args.workload} scene generator)",
        "config": {
            "workload": workload,
            "width": W, "height": H, "spp": spp, "max_depth": desc.maxBounces, "triangles": desc.numTris,
            "iteration_cap": "none (paths run to completion)", "paths_in_flight": args.max_paths,
        },
        "mrays_per_s": round(mrays, 3),
        "bvh_build_ms": round(bvh_build_ms, 2),
        "kernel_ms_per_step": {k: round(v / args.steps, 3) for k, v in totals.items()},
        "roofline": {
            "kernel": "k_trace<closest>",
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_unit": "bytes per launch (HBM, rocprofv3 FETCH_SIZE+WRITE_SIZE, calibrated)",
            "algorithmic_bytes_per_launch": round(bytes_per_ray * rays_c / max(launches_c, 1)),
            "bytes_per_ray": round(bytes_per_ray, 2),
            "nodes_per_ray": round(nodes_per_ray, 3),
            "tris_per_ray": round(tris_per_ray, 3),
            "avg_launch_ms": round(avg_launch_ms, 4),
            "rays_per_launch": round(rays_c / max(launches_c, 1), 1),
            "any_hit_gbs": round(any_gbs, 2),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def cpu_baseline(desc, W, H, target_s):
    """Oracle (the CPU restatement of libWR's kernel math) on the host cores."""
    from oracle import pyoracle as po

    threads = max(1, min(16, os.cpu_count() or 1))
    log(f"[bench] CPU baseline: oracle with host BVH, {threads} threads")
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
            "sample": f"{done_spp} spp of the same {W}x{H} frame (samples 0..{done_spp - 1}), full depth 16"}


if __name__ == "__main__":
    main()
