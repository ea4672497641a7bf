#!/usr/bin/env python3
"""Throughput benchmark of the MI355X wavefront path tracer.

Workload (one step = one full frame): the 1M-triangle random scene - 3125
frequency-4 icospheres (1,000,000 triangles) + 4 quad lights, Disney
materials, 1920x1080, 64 spp, max depth 8, NEE + MIS on - rendered through the
C ABI (srt_render) into a device-resident film.  Inputs (scene, BVH) are
resident in HBM before the timed region; the BVH build is timed separately.

Output: one JSON line with Mpaths/s (whole frame: raygen -> traverse -> shade
-> shadow -> finalize), per-request latency percentiles over the timed steps,
the roofline of the dominant kernel (algorithmic bytes from the node-visit and
triangle-test counts of an untimed counting pass over the same frame, divided
by that kernel's mean launch time from HIP events on its stream), and the
CPU baseline: the oracle (host execution of the reference integrator, float)
on a fixed subset of rows of the same frame, on all host cores, whose pixels
are also compared with the GPU frame.

Usage: python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Mpaths/s at equal RMSE vs FP64 oracle; p50/p99 render-request latency"

# algorithmic bytes per unit (SoA layout of srt_amd/csrc/srt_internal.h)
RAY_IN = 24          # origin + direction read by the traversal kernel
HIT_OUT = 16         # t, u, v, triangle written by the traversal kernel
NODE = 64            # one BVH2 node (two child boxes + refs)
TRI = 36             # three vertices of a tested triangle
SHADOW_IN = 44       # shadow item: origin, direction, tmax, contribution, slot
L_RMW = 24           # radiance accumulator read + write (3 floats)
QUEUE_ENTRY = 56     # ray + path state of one queue entry (read or write)
TRI_FETCH = 48       # hit triangle record read by shade
MATERIAL = 80        # material record read by shade


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-row-step", type=int, default=8, help="CPU baseline renders every k-th row")
    ap.add_argument("--path-pool", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the frame would shard by "
              "image tiles or spp slices with one gather/all-reduce, not implemented); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    from srt_amd import abi, scenes

    log("generating the 1M-triangle scene")
    s = scenes.spheres()
    R = s.render
    W, H, spp, depth = R["width"], R["height"], R["spp"], R["max_depth"]
    paths = W * H * spp
    ctx = abi.Context(0)
    if args.path_pool:
        ctx.set_option("path_pool", args.path_pool)
    h = ctx.upload(s)
    t0 = time.perf_counter()
    ctx.build_bvh(h)
    bvh_ms = (time.perf_counter() - t0) * 1e3
    info = ctx.bvh_info(h)
    log(f"BVH: {info['num_nodes']} nodes, depth {info['max_depth']}, {info['clusters']} clusters, {bvh_ms:.1f} ms")

    # untimed counting pass over the same frame (node visits / triangle tests)
    ctx.set_option("count_visits", 1)
    ctx.render_resident(h, s.camera, W, H, spp, depth, R["rr_depth"], 1, 1, R["seed"])
    visits = ctx.last_visits()
    segments, shadow_rays = ctx.last_counts()
    ctx.set_option("count_visits", 0)
    log(f"segments {segments}, shadow rays {shadow_rays}, visits {visits}")

    for _ in range(args.warmup):
        ctx.render_resident(h, s.camera, W, H, spp, depth, R["rr_depth"], 1, 1, R["seed"])
    ctx.set_option("profile", 1)
    ctx.stats_reset()
    lat = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t1 = time.perf_counter()
        ctx.render_resident(h, s.camera, W, H, spp, depth, R["rr_depth"], 1, 1, R["seed"])   # synchronizes
        lat.append((time.perf_counter() - t1) * 1e3)
    elapsed = time.perf_counter() - t_start
    ctx.set_option("profile", 0)
    st = ctx.stats()
    gpu_img = ctx.film(W * H)
    value = paths * args.steps / elapsed / 1e6
    log(f"{value:.1f} Mpaths/s, {elapsed / args.steps * 1e3:.1f} ms/frame; kernels {json.dumps(st)}")

    # algorithmic bytes per frame per kernel
    emitted = segments - paths     # queue entries written by shade
    bytes_step = {
        "trace": RAY_IN * segments + HIT_OUT * segments + NODE * visits["closest_nodes"] + TRI * visits["closest_tris"],
        "shadow": (SHADOW_IN + L_RMW) * shadow_rays + NODE * visits["shadow_nodes"] + TRI * visits["shadow_tris"],
        "shade": (QUEUE_ENTRY + HIT_OUT + TRI_FETCH + MATERIAL) * segments + QUEUE_ENTRY * emitted
                 + SHADOW_IN * shadow_rays,
    }
    kern = max((k for k in st if k in bytes_step), key=lambda k: st[k]["ms"])
    launches_per_step = st[kern]["launches"] / args.steps
    avg_ms = st[kern]["ms"] / st[kern]["launches"]
    bytes_launch = bytes_step[kern] / launches_per_step
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": kern, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "bytes_per_launch": int(bytes_launch), "avg_launch_ms": round(avg_ms, 4),
                "per_kernel_GBs": {k: round(bytes_step[k] / (st[k]["ms"] / args.steps * 1e-3) / 1e9, 1)
                                   for k in bytes_step if k in st and st[k]["ms"] > 0}}

    cpu = None
    if not args.no_cpu:
        from oracle import pyoracle as O   # CPU baseline leg only
        threads = min(os.cpu_count() or 1, 16)
        log(f"CPU baseline: oracle on every {args.cpu_row_step}th row, {threads} threads")
        osc = O.Scene(s)
        ob = O.Bvh(osc)
        img, secs = O.render(osc, s.camera, W, H, spp, depth, R["rr_depth"], 1, 1, R["seed"], bvh=ob,
                             threads=threads, row_begin=0, row_step=args.cpu_row_step)
        rows = list(range(0, H, args.cpu_row_step))
        cpu_paths = len(rows) * W * spp
        a = gpu_img.reshape(H, W, 4)[rows, :, :3]
        b = img.reshape(H, W, 4)[rows, :, :3]
        rel = np.abs(a - b) / np.maximum(np.maximum(np.abs(a), np.abs(b)), 1e-6)
        cpu = {"value": round(cpu_paths / secs / 1e6, 3), "unit": "Mpaths/s", "cores": threads, "kind": "port",
               "sample": f"{len(rows)} of {H} rows (every {args.cpu_row_step}th) of the same 1920x1080x64spp frame, "
                         f"{cpu_paths} paths, {secs:.1f} s",
               "gpu_vs_cpu_max_rel_diff": float(rel.max())}
        log(f"CPU {cpu['value']} Mpaths/s, GPU/CPU {value / cpu['value']:.1f}x, max rel diff {rel.max():.2e}")

    lat_sorted = sorted(lat)
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mpaths/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "spheres-1M: 3125 frequency-4 icospheres (1,000,000 triangles) + 4 quad lights, "
                               "Disney, 1920x1080, 64 spp, max depth 8, NEE+MIS",
                   "width": W, "height": H, "spp": spp, "max_depth": depth, "triangles": s.num_triangles,
                   "paths_per_step": paths, "parallelism": "single-gpu"},
        "latency_ms": {"p50": round(float(np.percentile(lat_sorted, 50)), 2),
                       "p99": round(float(np.percentile(lat_sorted, 99)), 2), "unit": "ms per frame request"},
        "segments_per_path": round(segments / paths, 3), "shadow_rays_per_path": round(shadow_rays / paths, 3),
        "bvh_build_ms": round(bvh_ms, 1),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
