#!/usr/bin/env python3
"""Benchmark of the MI355X path-tracing hot path.

Headline line (one JSON object on stdout): closest-hit traversal throughput of the
traversal microbenchmark -- 1M triangles (centroids uniform in the unit cube, vertex
offsets N(0, 0.01)), 16M rays (origins uniform in [-0.5, 1.5]^3, directions uniform on
the sphere), 4-wide SAH BVH, FP32 watertight test.  One step = one closest-hit pass
over the 16M rays, with rays, BVH and outputs resident in HBM.

Also reported (key "path_tracing"): the room scene (256 x 256, 16 spp, max depth 8,
RR from depth 3) rendered with the `pt` estimator -- M path-samples/s and Mrays/s; and
(key "adjoint") the inverse-rendering workload -- the room plus 2,000 boxes (96k
triangles), 512 x 512 x 32 spp, max depth 8, L2 loss against a target rendered with hidden
box reflectances -- forward and path-replay backward times and their ratio.

`roofline` is measured live: HIP events around every traversal launch on the launch
stream give the average kernel duration; algorithmic bytes per launch come from the
kernel's own counting run (node visits x 112 B + triangle tests x 48 B + 32 B ray +
16 B hit per ray).  `cpu_baseline` times the CPU restatement (FP32 4-wide QBVH with SSE,
OpenMP over the host cores) on the same rays.
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

METRIC = "M path-samples/s + Mrays/s on 1 MI355X at fixed RMSE; backward/forward ratio"
HBM_PEAK_GBS = 8000.0
NODE_BYTES = 112  # bytes of a 4-wide node the kernel loads (6 box planes x 4 + 4 child words)
TRI_BYTES = 48    # one triangle record
RAY_BYTES = 32    # o, d, tmin, tmax
HIT_BYTES = 16    # t, b0, b1, prim


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_info():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model


def host_threads() -> int:
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        n = min(n, int(env))
    return max(1, n)


def traffic_from_profile():
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if present."""
    p = ROOT / "profiles" / "traversal_closest_1Mtri_16Mray_pmc.json"
    if not p.exists():
        return None
    try:
        d = json.loads(p.read_text())
        return float(d["hbm_bytes_per_launch"])
    except Exception:
        return None


def run_traversal(args, torch, Device, scenes):
    t0 = time.time()
    sc = scenes.traversal_scene(args.triangles)
    rays = scenes.traversal_rays(args.rays)
    log(f"[bench] generated {sc.num_triangles} triangles and {len(rays)} rays in {time.time() - t0:.1f}s")
    dev = Device(0)
    dev.upload(sc)
    info = dev.build_bvh(max_leaf_size=args.leaf)
    log(f"[bench] BVH: {info.num_nodes} nodes, depth {info.max_depth}, {info.build_seconds:.2f}s, "
        f"{info.device_bytes / 2**20:.1f} MiB")
    cols = [torch.from_numpy(rays.o[:, k].copy()) for k in range(3)] + \
           [torch.from_numpy(rays.d[:, k].copy()) for k in range(3)] + \
           [torch.from_numpy(rays.tmin), torch.from_numpy(rays.tmax)]
    dcols = [c.to("cuda:0") for c in cols]
    n = len(rays)
    out = [torch.empty(n, dtype=torch.float32, device="cuda:0") for _ in range(3)] + \
          [torch.empty(n, dtype=torch.int32, device="cuda:0")]
    ray_ptrs = [c.data_ptr() for c in dcols]
    out_ptrs = [c.data_ptr() for c in out]
    torch.cuda.synchronize()

    for _ in range(args.warmup):
        dev.trace_closest_device(ray_ptrs, out_ptrs, n)
    dev.synchronize()
    torch.cuda.synchronize()

    # Timed region: K steps, host clock, bracketed by device synchronisation.
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dev.trace_closest_device(ray_ptrs, out_ptrs, n)
    dev.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    log(f"[bench] traversal: {args.steps} steps in {wall:.3f}s")

    # Kernel duration from HIP events on the launch stream (separate pass, same work).
    dev.enable_timing(True)
    dev.reset_timing()
    for _ in range(args.steps):
        dev.trace_closest_device(ray_ptrs, out_ptrs, n)
    dev.synchronize()
    k_ms, k_n = dev.kernel_timing(0)
    dev.enable_timing(False)
    avg_ms = k_ms / max(k_n, 1)

    # Algorithmic bytes per launch from the counting traversal.
    visits, tests = dev.trace_count(rays)
    alg_bytes = n * (RAY_BYTES + HIT_BYTES) + visits * NODE_BYTES + tests * TRI_BYTES
    prim = out[3].cpu().numpy()
    hit_rate = float((prim >= 0).mean())
    dev.close()
    res = {
        "ms_per_step": wall / args.steps * 1e3,
        "value": n * args.steps / wall / 1e6,
        "kernel_ms": avg_ms,
        "alg_bytes": alg_bytes,
        "node_visits_per_ray": visits / n,
        "tri_tests_per_ray": tests / n,
        "hit_rate": hit_rate,
        "bvh": info,
    }
    return res, sc, rays


def run_pt(args, Device, scenes):
    sc = scenes.room_scene(256, 256, 16)
    dev = Device(0)
    dev.upload(sc)
    dev.build_bvh()
    dev.render("pt", num_samples=sc.render.num_samples // 16)  # warm-up
    secs, rays = [], []
    for _ in range(args.pt_steps):
        _, st = dev.render("pt", film_precision="fp32")
        secs.append(st.seconds)
        rays.append(st.closest_rays)
    dev.close()
    s = sum(secs) / len(secs)
    return {
        "workload": "room: 5-quad box room + 2 rotated boxes + 0.25x0.25 light (17,12,4); 256x256, 16 spp "
                    "(1,048,576 samples), max_path_vertices 9, rr_depth 3; renderer pt; FP32 film",
        "M_path_samples_per_s": sc.render.num_samples / s / 1e6,
        "Mrays_per_s": (sum(rays) / len(rays)) / s / 1e6,
        "ms_per_render": s * 1e3,
        "rays_per_sample": (sum(rays) / len(rays)) / sc.render.num_samples,
    }, sc


def run_adjoint(args, Device, scenes):
    """Forward render and path-replay backward of the gradient workload (pt, no RR)."""
    import numpy as np

    sc = scenes.gradient_scene(512, 512, 32)
    boxes = np.asarray(sc.meta["box_materials"])
    albedo = np.array([m["R"] for m in sc.materials], np.float64)
    emission = np.array([l["Le"] for l in sc.lights], np.float64)
    hidden = albedo.copy()
    hidden[boxes] = scenes.hidden_box_albedo(len(boxes))
    dev = Device(0)
    dev.upload(sc)
    dev.build_bvh()
    dev.set_params(hidden, emission)
    target, _ = dev.render("pt", film_precision="fp32")
    dev.set_params(albedo, emission)
    fwd, bwd, res = [], [], []
    for _ in range(args.pt_steps):
        img, st = dev.render("pt", film_precision="fp32")
        fwd.append(st.seconds)
        d = (2.0 * (img - target)).astype(np.float32)
        ga, ge, info = dev.backward("pt", d)
        bwd.append(info["seconds"])
        res.append(info["replay_residual"])
    dev.close()
    f, b = sum(fwd) / len(fwd), sum(bwd) / len(bwd)
    return {
        "workload": "gradient: room + 2000 axis-aligned boxes (96,042 triangles), 4 ceiling lights; 512x512, "
                    "32 spp (8,388,608 samples), max_path_vertices 9, no RR; renderer pt; L2 loss vs a target "
                    "with hidden box reflectances U[0.1,0.9]; parameters: 2004 materials x RGB + 4 lights x RGB",
        "forward_ms": round(f * 1e3, 3),
        "backward_ms": round(b * 1e3, 3),
        "backward_over_forward": round(b / f, 3),
        "forward_M_path_samples_per_s": round(sc.render.num_samples / f / 1e6, 1),
        "replay_residual": max(res),
        "gradient_norm": float(np.linalg.norm(ga)),
    }


def cpu_baseline_traversal(sc, rays, sample_rays):
    from oracle import lmref

    threads = host_threads()
    o = lmref.Oracle(sc, fp64=False, accel="qbvh", approx_rsqrt=True)
    sub = rays.slice(0, min(sample_rays, len(rays)))
    t0 = time.perf_counter()
    o.trace_closest(sub, threads=threads)
    dt = time.perf_counter() - t0
    return {"value": len(sub) / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"first {len(sub)} of the 16M rays, closest hit, FP32 SSE 4-wide QBVH (the reference's "
                      f"default scene), OpenMP over {threads} threads of {cpu_info()}; BVH build excluded"}


def cpu_baseline_pt(sc, samples):
    from oracle import lmref

    threads = host_threads()
    o = lmref.Oracle(sc, fp64=False, accel="qbvh", approx_rsqrt=True)
    _, st = o.render("pt", num_samples=samples, rng="mt", threads=threads)
    return {"value": st["samples"] / st["seconds"] / 1e6, "unit": "M path-samples/s", "cores": threads,
            "kind": "port", "Mrays_per_s": st["rays"] / st["seconds"] / 1e6,
            "sample": f"{samples} path samples of the room scene (pt, mt scheduler, 100 samples per block, "
                      f"std::mt19937 streams), {threads} threads of {cpu_info()}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--triangles", type=int, default=1_000_000)
    ap.add_argument("--rays", type=int, default=16_000_000)
    ap.add_argument("--leaf", type=int, default=4)
    ap.add_argument("--pt-steps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-pt", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--cpu-sample-rays", type=int, default=16_000_000)
    ap.add_argument("--cpu-sample-paths", type=int, default=32 * 1_048_576)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (profiler byte-counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 3}), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              f"(the path shards by sample index; see DESIGN.md).", file=sys.stderr)
        return 2
    rank = int(os.environ.get("RANK", "0"))

    import torch

    from lightmetrica_amd import scenes
    from lightmetrica_amd.gpu import Device

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 1
    res, sc, rays = run_traversal(args, torch, Device, scenes)
    achieved = res["alg_bytes"] / (res["kernel_ms"] * 1e-3) / 1e9
    traffic = traffic_from_profile()
    line = {
        "metric": METRIC,
        "value": round(res["value"], 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": f"traversal microbench: {args.triangles} triangles (centroids U[0,1]^3, vertices + "
                        f"N(0,0.01)), {args.rays} rays (origins U[-0.5,1.5]^3, directions uniform on the sphere), "
                        f"closest hit, 4-wide SAH BVH (max leaf {args.leaf}), FP32 watertight test",
            "triangles": args.triangles,
            "rays": args.rays,
            "hit_rate": round(res["hit_rate"], 4),
            "node_visits_per_ray": round(res["node_visits_per_ray"], 2),
            "tri_tests_per_ray": round(res["tri_tests_per_ray"], 2),
            "bvh_nodes": res["bvh"].num_nodes,
            "bvh_depth": res["bvh"].max_depth,
            "bvh_build_s": round(res["bvh"].build_seconds, 2),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": "k_trace_closest",
            "kernel_ms": round(res["kernel_ms"], 4),
            "alg_bytes_per_launch": int(res["alg_bytes"]),
        },
    }
    if not args.no_pt:
        pt, room = run_pt(args, Device, scenes)
        line["path_tracing"] = pt
    if not args.no_adjoint:
        line["adjoint"] = run_adjoint(args, Device, scenes)
    if not args.no_cpu_baseline and rank == 0:
        line["cpu_baseline"] = cpu_baseline_traversal(sc, rays, args.cpu_sample_rays)
        if not args.no_pt:
            line["path_tracing"]["cpu_baseline"] = cpu_baseline_pt(room, args.cpu_sample_paths)
            line["path_tracing"]["speedup_vs_cpu"] = round(
                line["path_tracing"]["M_path_samples_per_s"] / line["path_tracing"]["cpu_baseline"]["value"], 1)
        line["speedup_vs_cpu"] = round(line["value"] / line["cpu_baseline"]["value"], 1)
    if rank == 0:
        print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
