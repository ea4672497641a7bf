#!/usr/bin/env python3
"""Throughput benchmark of the MI355X wavefront path tracer.

Workload (one step = one full frame): the 1M-triangle random scene - 3125
frequency-4 icospheres (1,000,000 triangles) + 4 quad lights, Disney
materials, 1920x1080, 64 spp, max depth 8, NEE + MIS on - rendered through the
C ABI (srt_render) into a device-resident film.  Inputs (scene, BVH) are
resident in HBM before the timed region; the BVH build is timed separately.

Output: one JSON line with Mpaths/s (whole frame: traverse -> shade -> shadow
per bounce, the first bounce computing its camera rays, -> finalize), per-request latency percentiles over the timed steps,
the roofline of the dominant kernel (algorithmic bytes from the node-visit and
triangle-test counts of an untimed counting pass over the same frame, divided
by that kernel's mean launch time from HIP events on its stream), and the
CPU baseline: the oracle (host execution of the reference integrator, float)
on a fixed subset of rows of the same frame, on all host cores, whose pixels
are also compared with the GPU frame.

HBM traffic (roofline.traffic) comes from two rocprofv3 counter runs of this
same command (FETCH_SIZE, then WRITE_SIZE, per dispatch), each scaled by what
the counter reads for a copy kernel of known size (--mode copy), condensed by
--pmc-summary into profiles/trace_spheres1m_hbm_traffic.md, which this script
reads when present.

Usage: python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--no-cpu]
       python bench.py --mode skin|mlp|serve|passes (skinned animation; radiance MLP at equal time; serving;
                                    Simple / LightTracer / Denoise pass throughput)
       python bench.py --mode copy          (calibration kernel for the counters)
       python bench.py --pmc-summary FETCH_DIR WRITE_DIR COPY_FETCH_DIR COPY_WRITE_DIR
       python bench.py --kstats-summary DIR  (rocprofv3 --kernel-trace --stats output -> profiles/)
       python bench.py --mlp-summary KS_DIR PMC_DIR  (MLP kernel times + MFMA counters -> profiles/)
       python bench.py --valu-summary DIR  (VALU issue counters of the frame kernels -> profiles/)
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# MI355X's published FP32 vector peak, 157.3 TFLOP/s, as lane operations (an FMA
# is 2 FLOP): 256 CUs x 4 SIMDs x 32 lanes per clock x 2.4 GHz = 78.6 T/s
VALU_PEAK_TLANE_OPS = 256 * 4 * 32 * 2.4e9 / 1e12
MFMA_BF16_PEAK_TFLOPS = 2500.0   # dense BF16 MFMA, MI355X
MLP_TRAIN_FLOP = 99072           # forward + backward + weight gradients per training sample (16,512 MACs x 6)
MLP_FWD_FLOP = 33024             # forward per query (16,512 MACs x 2)
METRIC = "Mpaths/s at equal RMSE vs FP64 oracle; p50/p99 render-request latency"

# algorithmic bytes per unit (SoA layout of srt_amd/csrc/srt_internal.h)
RAY_IN = 24          # origin + direction read by the traversal kernel (bounces after the first, whose
                     # camera rays are computed from the slot)
HIT_OUT = 12         # hit record: barycentrics u, v and the sorted triangle
NODE = 112           # one 4-wide node visit: 6 plane vectors (near/far x, y, z of 4 children) + 4 refs
TRI = 36             # three vertices of a tested triangle
SHADOW_IN = 44       # shadow item: origin, direction, tmax, contribution, slot
L_RMW = 24           # radiance accumulator read + write (3 floats)
QUEUE_ENTRY = 56     # ray + path state of one queue entry (read or write)
TRI_FETCH = 48       # hit triangle record read by shade
MATERIAL = 80        # material record read by shade


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "trace_spheres1m_hbm_traffic.md")
VALU_FILE = os.path.join(ROOT, "profiles", "trace_spheres1m_valu_issue.md")
MI355X_SIMDS, MI355X_XCDS = 256 * 4, 8
COPY_ELEMS = 1 << 28          # calibration copy: 1 GiB read + 1 GiB written (float32)
# kernels of the frame, as rocprofv3 names them
# the timed (not counting) 4-wide kernels of a flat scene; the first bounce's
# closest hits run in k_trace_cam (k_trace_packet with the option trace_packet)
KERNEL_KEYS = {"trace": ("k_trace<false, true, false>", "k_trace_cam<false, true, false>", "k_trace_packet("),
               "shadow": ("k_shadow<false, true, false>",),
               "shade": ("k_shade<", "k_shade(")}


def kernel_of(name):
    """The frame kernel (trace / shadow / shade) a rocprofv3 kernel name belongs to, or None."""
    for k, keys in KERNEL_KEYS.items():
        if any(key in name for key in keys):
            return k
    return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def copy_mode():
    """Calibration kernel for the HBM byte counters: an elementwise negation of
    2^28 float32 values (1 GiB read, 1 GiB written), run 3 times."""
    import torch
    x = torch.ones(COPY_ELEMS, device="cuda", dtype=torch.float32)
    y = torch.empty_like(x)
    for _ in range(3):
        torch.neg(x, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": COPY_ELEMS * 4, "copy_bytes_written": COPY_ELEMS * 4}), flush=True)


def _pmc_rows(d, counter):
    """(kernel name, value) per dispatch from a rocprofv3 --pmc output folder
    (rocpd SQLite database, or counter_collection.csv with -f csv)."""
    import csv
    import glob
    import sqlite3
    rows = []
    dbs = glob.glob(os.path.join(d, "**", "*results.db"), recursive=True)
    for f in dbs:
        con = sqlite3.connect(f)
        rows += [(n, float(v)) for n, v in
                 con.execute("select kernel_name, value from counters_collection where counter_name = ?", (counter,))]
        con.close()
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                if r["Counter_Name"] == counter:
                    rows.append((r["Kernel_Name"], float(r["Counter_Value"])))
    if not rows:
        raise SystemExit(f"no {counter} values under {d}")
    return rows


def pmc_summary(fetch_dir, write_dir, copy_fetch_dir, copy_write_dir):
    """Condense the counter runs into the traffic summary (KiB per dispatch ->
    calibrated bytes per launch for each frame kernel)."""
    def per_kernel(d, counter):
        acc = {}
        for name, v in _pmc_rows(d, counter):
            k = kernel_of(name)
            if k:
                acc.setdefault(k, []).append(v)
        return {k: (sum(v) / len(v), len(v)) for k, v in acc.items()}

    def copy_kib(d, counter):
        vals = [v for name, v in _pmc_rows(d, counter) if "neg" in name.lower()]
        if not vals:
            raise SystemExit(f"no negation kernel in {d}")
        return sum(vals) / len(vals)

    known_kib = COPY_ELEMS * 4 / 1024.0
    cf, cw = copy_kib(copy_fetch_dir, "FETCH_SIZE"), copy_kib(copy_write_dir, "WRITE_SIZE")
    ff, fw = known_kib / cf, known_kib / cw
    fetch, write = per_kernel(fetch_dir, "FETCH_SIZE"), per_kernel(write_dir, "WRITE_SIZE")
    lines = ["# HBM traffic per launch of the frame kernels (1M-triangle frame, 1920x1080x64spp, depth 8)", "",
             "Commands (from the repository root, then this summary):", "",
             "    rocprofv3 --pmc FETCH_SIZE -d gpurun_out/pmc_FETCH_SIZE -o run -- python3 bench.py --steps 1 --warmup 0 --no-cpu --no-serve",
             "    rocprofv3 --pmc WRITE_SIZE -d gpurun_out/pmc_WRITE_SIZE -o run -- python3 bench.py --steps 1 --warmup 0 --no-cpu --no-serve",
             "    rocprofv3 --pmc FETCH_SIZE -d gpurun_out/pmc_copy_FETCH_SIZE -o run -- python3 bench.py --mode copy",
             "    rocprofv3 --pmc WRITE_SIZE -d gpurun_out/pmc_copy_WRITE_SIZE -o run -- python3 bench.py --mode copy",
             "    python3 bench.py --pmc-summary gpurun_out/pmc_FETCH_SIZE gpurun_out/pmc_WRITE_SIZE "
             "gpurun_out/pmc_copy_FETCH_SIZE gpurun_out/pmc_copy_WRITE_SIZE", "",
             "Each rocprofv3 run: cd /tmp && export TMPDIR=/tmp, bench.py by its path in the repository.",
             "The frame runs hold one counting pass (k_trace<true, true>/k_shadow<true, true>, not counted here; its k_shade",
             "dispatches are) and one frame; trace is k_trace_cam (first bounce) and k_trace (later bounces);",
             "FETCH_SIZE/WRITE_SIZE are the L2-to-fabric bytes (MALL or HBM).", "",
             f"Calibration: copy kernel of {COPY_ELEMS * 4} B read + {COPY_ELEMS * 4} B written: "
             f"FETCH_SIZE {cf:.1f} KiB, WRITE_SIZE {cw:.1f} KiB per dispatch -> factors fetch {ff:.4f}, "
             f"write {fw:.4f}", "",
             "| kernel | dispatches | FETCH_SIZE KiB/dispatch | WRITE_SIZE KiB/dispatch | calibrated bytes/launch |",
             "|---|---|---|---|---|"]
    for k in KERNEL_KEYS:
        if k in fetch and k in write:
            b = (fetch[k][0] * ff + write[k][0] * fw) * 1024.0
            lines.append(f"| {k} | {fetch[k][1]} | {fetch[k][0]:.1f} | {write[k][0]:.1f} | {b:.0f} |")
    lines.append("")
    for k in KERNEL_KEYS:
        if k in fetch and k in write:
            b = (fetch[k][0] * ff + write[k][0] * fw) * 1024.0
            lines.append(f"traffic_bytes_per_launch {k} {b:.0f}")
    text = "\n".join(lines) + "\n"
    os.makedirs(os.path.dirname(TRAFFIC_FILE), exist_ok=True)
    with open(TRAFFIC_FILE, "w") as fh:
        fh.write(text)
    print(text)


def valu_summary(d):
    """VALU issue of the frame kernels from one rocprofv3 --pmc pass: the SIMDs'
    busy fraction 4 SQ_ACTIVE_INST_VALU / (SIMDs x GRBM_GUI_ACTIVE / XCDs)
    (a wave64 VALU instruction holds its SIMD for one quad-cycle; GRBM_GUI_ACTIVE
    sums the cycles of the 8 XCDs) and the active lanes per VALU instruction,
    SQ_THREAD_CYCLES_VALU / (64 SQ_ACTIVE_INST_VALU)."""
    names = ["SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_THREAD_CYCLES_VALU", "SQ_WAVE_CYCLES", "SQ_WAIT_ANY",
             "GRBM_GUI_ACTIVE"]
    per = {}
    for cn in names:
        for name, v in _pmc_rows(d, cn):
            k = kernel_of(name)
            if k:
                per.setdefault(k, {}).setdefault(cn, []).append(v)
    lines = ["# VALU issue of the frame kernels (1M-triangle frame, 1920x1080x64spp, depth 8)", "",
             "Commands (from the repository root, then this summary):", "",
             "    rocprofv3 --pmc " + " ".join(names) + " -d gpurun_out/pmc_valu -o run -- python3 bench.py "
             "--steps 1 --warmup 0 --no-cpu --no-serve",
             "    python3 bench.py --valu-summary gpurun_out/pmc_valu", "",
             "The rocprofv3 run: cd /tmp && export TMPDIR=/tmp, bench.py by its path in the repository; the counting",
             "pass's k_trace<true, true>/k_shadow<true, true> are not counted here (its k_shade dispatches are); trace",
             "is k_trace_cam (first bounce) and k_trace (later bounces).",
             f"busy = 4 SQ_ACTIVE_INST_VALU / ({MI355X_SIMDS} SIMDs x GRBM_GUI_ACTIVE / {MI355X_XCDS}), summed over the "
             "kernel's dispatches; lanes = SQ_THREAD_CYCLES_VALU / (64 SQ_ACTIVE_INST_VALU); wait = SQ_WAIT_ANY / "
             "SQ_WAVE_CYCLES.", "",
             "| kernel | dispatches | VALU instructions | SIMD VALU busy | active lanes per VALU instruction | wave cycles waiting |",
             "|---|---|---|---|---|---|"]
    out = {}
    for k in KERNEL_KEYS:
        c = per.get(k)
        if not c or any(n not in c for n in names):
            continue
        tot = {n: sum(c[n]) for n in names}
        busy = 4.0 * tot["SQ_ACTIVE_INST_VALU"] / (MI355X_SIMDS * tot["GRBM_GUI_ACTIVE"] / MI355X_XCDS)
        lanes = tot["SQ_THREAD_CYCLES_VALU"] / (64.0 * tot["SQ_ACTIVE_INST_VALU"])
        wait = tot["SQ_WAIT_ANY"] / tot["SQ_WAVE_CYCLES"]
        out[k] = (busy, lanes)
        lines.append(f"| {k} | {len(c['SQ_INSTS_VALU'])} | {tot['SQ_INSTS_VALU']:.4g} | {busy:.3f} | {lanes:.3f} | "
                     f"{wait:.3f} |")
    lines.append("")
    for k, (busy, lanes) in out.items():
        lines.append(f"valu_issue {k} {busy:.4f} {lanes:.4f}")
    text = "\n".join(lines) + "\n"
    with open(VALU_FILE, "w") as fh:
        fh.write(text)
    print(text)


def measured_valu_counts(kernel):
    """(wave64 VALU instructions per launch, active lanes per instruction, SIMD
    busy fraction, source file) of `kernel` from the committed counter summary."""
    if not os.path.exists(VALU_FILE):
        return None
    insts = None
    with open(VALU_FILE) as fh:
        for line in fh:
            c = [x.strip() for x in line.strip().strip("|").split("|")]
            if len(c) == 6 and c[0] == kernel:
                try:
                    insts = float(c[2]) / float(c[1])
                except ValueError:
                    pass
    v = measured_valu(kernel)
    if insts is None or v is None:
        return None
    return insts, v[1], v[0], os.path.relpath(VALU_FILE, ROOT)


def measured_valu(kernel):
    """(SIMD VALU busy fraction, active lanes per VALU instruction) of `kernel`
    from the committed counter summary."""
    if not os.path.exists(VALU_FILE):
        return None
    with open(VALU_FILE) as fh:
        for line in fh:
            p = line.split()
            if len(p) == 4 and p[0] == "valu_issue" and p[1] == kernel:
                return float(p[2]), float(p[3])
    return None


KSTATS_FILE = os.path.join(ROOT, "profiles", "frame_spheres1m_kernel_stats.md")


PASSES_KSTATS_FILE = os.path.join(ROOT, "profiles", "passes_spheres1m_kernel_stats.md")
SKIN_KSTATS_FILE = os.path.join(ROOT, "profiles", "skin_tube200k_kernel_stats.md")
INST_KSTATS_FILE = os.path.join(ROOT, "profiles", "frame_spheres1m_instanced_kernel_stats.md")


def kstats_summary(d, mode="frame"):
    """Per-kernel time summary of a `rocprofv3 --kernel-trace --stats` run of
    this script (rocpd database or kernel_stats.csv); mode "frame" (the default
    bench) or "passes" (--mode passes)."""
    import csv
    import glob
    import sqlite3
    rows = []
    for f in glob.glob(os.path.join(d, "**", "*results.db"), recursive=True):
        con = sqlite3.connect(f)
        # the database view holds microseconds
        rows += [(n, int(c), float(t) * 1e3, float(a) * 1e3) for n, c, t, a in
                 con.execute("select name, total_calls, total_duration, average from top_kernels")]
        con.close()
    for f in glob.glob(os.path.join(d, "**", "*kernel_stats.csv"), recursive=True):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                rows.append((r["Name"], int(r["Calls"]), float(r["TotalDurationNs"]), float(r["AverageNs"])))
    if not rows:
        raise SystemExit(f"no kernel statistics under {d}")
    rows.sort(key=lambda r: -r[2])
    total = sum(r[2] for r in rows)
    if mode == "skin":
        lines = ["# Kernel times: skinned 200k-triangle tube, 60 frames of skin + refit + 1280x720x16spp depth 6", "",
                 "Command (from the repository root; TMPDIR=/tmp):", "",
                 "    rocprofv3 --kernel-trace --stats -d gpurun_out/prof_skin -o run -- python3 bench.py --mode skin "
                 "--warmup 1",
                 "    python3 bench.py --kstats-summary gpurun_out/prof_skin --kstats-mode skin", "",
                 "The run builds the BVH once, then renders 1 warmup frame and 60 timed frames; every frame",
                 "skins the tube (k_skin), refits the binary tree level by level (k_refit_level) and the 4-wide",
                 "tree (k_refit4), then renders (k_trace_cam / k_trace, k_shade, k_shadow, k_finalize).", "",
                 "| kernel | calls | total ms | average us | share |", "|---|---|---|---|---|"]
        out_file = SKIN_KSTATS_FILE
    elif mode == "passes":
        lines = ["# Kernel times: Simple, LightTracer and Denoise passes on the 1M-triangle scene (1920x1080)", "",
                 "Command (from the repository root; TMPDIR=/tmp):", "",
                 "    rocprofv3 --kernel-trace --stats -d gpurun_out/prof_passes -o run -- python3 bench.py --mode passes "
                 "--steps 3 --warmup 1",
                 "    python3 bench.py --kstats-summary gpurun_out/prof_passes --kstats-mode passes", "",
                 "Each pass runs 1 warmup + 3 timed calls: Simple Depth and Normal at 4 samples per pixel",
                 "(k_simple_trace, k_simple_accum), LightTracer at 1 particle per pixel and depth 8 (k_lt_raygen, k_trace,",
                 "k_lt_shade, k_lt_shadow), Denoise (k_denoise_guide, 5 x k_atrous); the BVH build runs once.", "",
                 "| kernel | calls | total ms | average us | share |", "|---|---|---|---|---|"]
        out_file = PASSES_KSTATS_FILE
    elif mode == "instanced":
        lines = ["# Kernel times: 1M-triangle frame in two-level form (1920x1080x64spp, depth 8)", "",
                 "Command (from the repository root; TMPDIR=/tmp):", "",
                 "    rocprofv3 --kernel-trace --stats -d gpurun_out/prof_inst -o run -- python3 bench.py --steps 2 "
                 "--warmup 1 --no-cpu --no-serve --scene spheres_instanced",
                 "    python3 bench.py --kstats-summary gpurun_out/prof_inst --kstats-mode instanced", "",
                 "16 icosphere meshes (320 triangles each) under 3125 instances plus 4 light instances; the run",
                 "renders 6 frames (the counting pass, 1 warmup, 2 timed, 2 copied to host memory) with the",
                 "instanced traversal kernels (k_trace<.., .., true>, k_shadow<.., .., true>); the two-level tree",
                 "is built twice (cold, then a rebuild), its per-mesh trees by the same kernels as the flat build.", "",
                 "| kernel | calls | total ms | average us | share |", "|---|---|---|---|---|"]
        out_file = INST_KSTATS_FILE
    else:
        lines = ["# Kernel times: 1M-triangle frame (1920x1080x64spp, depth 8)", "",
                 "Command (from the repository root; TMPDIR=/tmp):", "",
                 "    rocprofv3 --kernel-trace --stats -d gpurun_out/prof_ks -o run -- python3 bench.py --steps 2 --warmup 1 --no-cpu --no-serve",
                 "    python3 bench.py --kstats-summary gpurun_out/prof_ks", "",
                 "The run renders 6 frames: the untimed counting pass (k_trace_cam<true, true>, k_trace<true, true>,",
                 "k_shadow<true, true>), 1 warmup, 2 timed frames and 2 frames copied to host memory; the first bounce's",
                 "closest hits run in k_trace_cam (camera rays computed from the slot), later ones in k_trace; the BVH",
                 "build runs twice (cold, then a rebuild): the",
                 "top-level SAH levels (k_wide_* for levels of <= 64 nodes, k_top_split, k_top_split_wave for levels of",
                 "> 1024 nodes, k_top_link / k_link_*), k_bottom_up_level, sorts, 4-wide collapse.", "",
                 "| kernel | calls | total ms | average us | share |", "|---|---|---|---|---|"]
        out_file = KSTATS_FILE
    for n, c, t, a in rows[:16]:
        short = n.split("(")[0].replace("void ", "")
        if len(short) > 70:
            short = short[:67] + "..."
        lines.append(f"| `{short}` | {c} | {t / 1e6:.3f} | {a / 1e3:.1f} | {100 * t / total:.1f}% |")
    text = "\n".join(lines) + "\n"
    with open(out_file, "w") as fh:
        fh.write(text)
    print(text)


MLP_PROFILE_FILE = os.path.join(ROOT, "profiles", "mlp_train_cornell_mfma.md")
MLP_PMC = ["SQ_INSTS_VALU_MFMA_MOPS_BF16"]


def mlp_summary(ks_dir, pmc_dir):
    """MFMA summary of the radiance-MLP kernels: kernel times from a
    `rocprofv3 --kernel-trace --stats` run and BF16 MFMA operation counts from
    a `rocprofv3 --pmc` run of `bench.py --mode mlp --mlp-frames 4`."""
    import glob
    import sqlite3
    times = {}
    for f in glob.glob(os.path.join(ks_dir, "**", "*results.db"), recursive=True):
        con = sqlite3.connect(f)
        for n, c, t, a in con.execute("select name, total_calls, total_duration, average from top_kernels"):
            times[n] = (int(c), float(t) * 1e3, float(a) * 1e3)     # ns
        con.close()
    if not times:
        raise SystemExit(f"no kernel statistics under {ks_dir}")
    mops = {}
    for n, v in _pmc_rows(pmc_dir, MLP_PMC[0]):
        mops.setdefault(n, []).append(v)
    mlp_names = sorted((n for n in times if "mlp" in n), key=lambda n: -times[n][1])
    lines = ["# Radiance-MLP kernels: time and MFMA counter (Cornell 256x256 training frames)", "",
             "Commands (from the repository root; TMPDIR=/tmp):", "",
             "    rocprofv3 --kernel-trace --stats -d gpurun_out/mlp_ks -o run -- python3 bench.py --mode mlp --mlp-frames 4",
             "    rocprofv3 --pmc " + " ".join(MLP_PMC) + " -d gpurun_out/mlp_pmc -o run -- python3 bench.py --mode mlp "
             "--mlp-frames 4",
             "    python3 bench.py --mlp-summary gpurun_out/mlp_ks gpurun_out/mlp_pmc", "",
             "The run trains on 6 frames of about 120k vertex-2 samples each; srt_mlp_train_frame splits a",
             "frame's samples into 8 mini-batches, one Adam step each, so a k_mlp_train dispatch trains on about",
             "15k samples.  It renders the MLP form twice at 16 spp and three times at 256 spp (k_mlp_infer, one",
             "launch per render).  Counter GFLOP = SQ_INSTS_VALU_MFMA_MOPS_BF16 x 512 per dispatch, as read: for",
             "k_mlp_infer the FLOP of its MFMA instructions (36 per 32 queries), for k_mlp_train those of its",
             "15k samples (98 MFMA instructions per 32 samples: 1.59 GFLOP / 106k counter FLOP per sample).",
             "The rates in `bench.py --mode mlp` are algorithmic (99,072 FLOP per training sample, 33,024 per",
             "query) over HIP-event time.", "",
             "| kernel | calls | average us | counter GFLOP/dispatch |",
             "|---|---|---|---|"]
    for n in mlp_names:
        c, tot, avg = times[n]
        short = n.split("(")[0].replace("void ", "")
        flop = f"{sum(mops[n]) / len(mops[n]) * 512 / 1e9:.3f}" if n in mops else "-"
        lines.append(f"| `{short}` | {c} | {avg / 1e3:.1f} | {flop} |")
    text = "\n".join(lines) + "\n"
    with open(MLP_PROFILE_FILE, "w") as fh:
        fh.write(text)
    print(text)


def _stats_ms(st, name, steps):
    return round(st[name]["ms"] / steps, 3) if name in st else None


DENOISE_BYTES_PER_PIXEL = 32 + 5 * 48   # guide: 16 in + 16 out; each of 5 passes: colour + guide in, colour out


def passes_mode(args):
    """The other passes on the 1M-triangle scene at 1920x1080, output resident
    in device memory: Simple (Depth and Normal, 4 samples per pixel = camera
    rays traced), LightTracer (1 particle per pixel, depth 8) and Denoise
    (five a-trous passes over the float4 image)."""
    from srt_amd import abi, scenes
    s = scenes.spheres()
    R = s.render
    W, H = R["width"], R["height"]
    npix = W * H
    ctx = abi.Context(0)
    h = ctx.upload(s)
    ctx.build_bvh(h)
    img = ctx.device_alloc(16 * npix)
    reps = max(1, args.steps)
    out = {"metric": "throughput of the Simple, LightTracer and Denoise passes", "n_gpus": 1, "dtype": "f32",
           "data": "synthetic", "steps": reps,
           "config": {"workload": "spheres-1M: 3125 frequency-4 icospheres (1,000,000 triangles) + 4 quad lights, "
                                  "1920x1080", "width": W, "height": H}}

    def timed(fn, key):
        for _ in range(max(1, args.warmup)):
            fn()
        ctx.set_option("profile", 1)
        ctx.stats_reset()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        wall = (time.perf_counter() - t0) / reps
        ctx.set_option("profile", 0)
        st = ctx.stats()
        return wall, st[key]["ms"] / st[key]["launches"] if key in st and st[key]["launches"] else None

    spp = 4
    for t in ("Depth", "Normal"):
        wall, kms = timed(lambda: ctx.render_simple(h, s.camera, W, H, type=t, spp=spp, seed=1, device_ptr=img),
                          "simple")
        out[f"simple_{t.lower()}"] = {"spp": spp, "ms": round(wall * 1e3, 3), "gpu_ms": round(kms, 3),
                                      "Mrays_s": round(npix * spp / (kms * 1e-3) / 1e6, 1)}
    wall, kms = timed(lambda: ctx.render_light(h, s.camera, W, H, spp=1, max_depth=8, rr_depth=4, seed=2,
                                               device_ptr=img), "light_tracer")
    out["light_tracer"] = {"spp": 1, "max_depth": 8, "ms": round(wall * 1e3, 3), "gpu_ms": round(kms, 3),
                           "Mpaths_s": round(npix / (kms * 1e-3) / 1e6, 1)}
    rng = np.random.default_rng(3)
    noisy = np.ones((npix, 4), np.float32)
    noisy[:, :3] = rng.gamma(2.0, 0.5, size=(npix, 3)).astype(np.float32)
    ctx.upload_bytes(img, noisy)
    wall, kms = timed(lambda: ctx.denoise_device(img, W, H), "denoise")
    achieved = DENOISE_BYTES_PER_PIXEL * npix / (kms * 1e-3) / 1e9
    out["denoise"] = {"ms": round(wall * 1e3, 3), "gpu_ms": round(kms, 3),
                      "Mpixels_s": round(npix / (kms * 1e-3) / 1e6, 1),
                      "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                                   "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                                   "bytes_per_pixel": DENOISE_BYTES_PER_PIXEL}}
    ctx.device_free(img)
    ctx.release(h)
    ctx.close()
    print(json.dumps(out), flush=True)


def skin_mode(args):
    """Skinned 200k-triangle tube (64-bone chain, 4 bones per vertex): per frame
    linear blend skinning + BVH refit + a 1280x720, 16 spp, depth 6 render;
    60 frames (one bone-animation period)."""
    from srt_amd import abi, scenes
    s = scenes.tube()
    R = s.render
    W, H, spp, depth = R["width"], R["height"], R["spp"], R["max_depth"]
    frames = s.skin["frames"]
    mats = [s.skin["matrices"](f) for f in range(frames)]
    ctx = abi.Context(0)
    h = ctx.upload(s)
    t0 = time.perf_counter()
    ctx.build_bvh(h)
    bvh_ms = (time.perf_counter() - t0) * 1e3
    ctx.set_skin(h, s.skin["bone_ids"], s.skin["bone_weights"])
    for f in range(min(args.warmup, frames)):
        ctx.skin_refit(h, mats[f])
        ctx.render_resident(h, s.camera, W, H, spp, depth, R["rr_depth"], 1, 1, R["seed"] + f)
    ctx.set_option("profile", 1)
    ctx.stats_reset()
    skin_ms, render_ms = [], []
    t_start = time.perf_counter()
    for f in range(frames):
        t1 = time.perf_counter()
        ctx.skin_refit(h, mats[f])                       # synchronizes
        t2 = time.perf_counter()
        ctx.render_resident(h, s.camera, W, H, spp, depth, R["rr_depth"], 1, 1, R["seed"] + f)
        t3 = time.perf_counter()
        skin_ms.append((t2 - t1) * 1e3)
        render_ms.append((t3 - t2) * 1e3)
    elapsed = time.perf_counter() - t_start
    ctx.set_option("profile", 0)
    st = ctx.stats()
    paths = W * H * spp * frames
    nv = s.positions.shape[0]
    # k_skin algorithmic bytes: per vertex 4 ids + 4 weights (32 B) read, original
    # position/normal/tangent read and skinned ones written (12 B each)
    skin_bytes = nv * (32 + 3 * 12 + 3 * 12)
    out = {"metric": "Mpaths/s, skinned animation (skin + refit + render per frame)",
           "value": round(paths / elapsed / 1e6, 2), "unit": "Mpaths/s", "n_gpus": 1, "frames": frames,
           "ms_per_frame": round(elapsed / frames * 1e3, 2), "higher_is_better": True, "dtype": "f32",
           "data": "synthetic",
           "config": {"workload": "tube-200k: 1000x100x2 triangles, 64-bone chain, 4 bones/vertex (Dirichlet "
                                  "weights), floor + light; 1280x720, 16 spp, depth 6, 60 frames",
                      "triangles": s.num_triangles, "vertices": nv, "bones": s.skin["num_bones"]},
           "skin_refit_ms": {"mean": round(float(np.mean(skin_ms)), 3), "max": round(float(np.max(skin_ms)), 3)},
           "render_ms": {"mean": round(float(np.mean(render_ms)), 2), "max": round(float(np.max(render_ms)), 2)},
           "kernel_ms_per_frame": {k: _stats_ms(st, k, frames) for k in ("skin", "refit", "trace", "shade", "shadow")},
           "skin_GBs": (round(skin_bytes / (st["skin"]["ms"] / frames * 1e-3) / 1e9, 1)
                        if "skin" in st and st["skin"]["ms"] > 0 else None),
           "bvh_build_ms": round(bvh_ms, 1)}
    ctx.close()
    print(json.dumps(out), flush=True)


def anim_mode(args):
    """Node animation of the 1M-triangle sphere field: every sphere turns about
    its own centre and bobs, per frame.  Flat form: each sphere's vertices are
    one rigid bone (linear blend skinning with weight 1), then the whole BVH
    is refit.  Two-level form: each sphere is its own mesh under one instance,
    and a frame sets the 3129 instance transforms and refits the top-level
    tree (the reference's node animation: instance transforms, then an IAS
    update).  Each frame is then rendered at the full configuration."""
    from srt_amd import abi, scenes
    s = scenes.spheres()
    R = s.render
    W, H, spp, depth = R["width"], R["height"], R["spp"], R["max_depth"]
    frames = max(1, args.steps)
    nsph = 3125
    rng = np.random.default_rng(7)
    axes = rng.normal(size=(nsph, 3))
    phase = rng.uniform(0.0, 2.0 * math.pi, size=nsph)
    first = s.mesh_first
    centres = np.stack([s.positions[s.indices[first[i]:first[i + 1]].ravel()].mean(0) for i in range(nsph)])

    def motion(f):
        xf = np.zeros((nsph + 4, 3, 4), np.float32)
        xf[:, :, :3] = np.eye(3, dtype=np.float32)
        for i in range(nsph):
            Rm = scenes._rotation(axes[i], 0.05 * f)
            t = centres[i] - Rm @ centres[i] + np.array([0.0, 0.5 * math.sin(0.2 * f + phase[i]), 0.0])
            xf[i, :, :3] = Rm
            xf[i, :, 3] = t
        return xf

    moves = [motion(f + 1) for f in range(frames)]
    out = {"metric": "node animation: ms per frame (update + render), 1M-triangle sphere field, every sphere moving",
           "frames": frames, "higher_is_better": False, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "spheres-1M (3125 spheres, 1,000,008 triangles) with every sphere rigidly animated; "
                                  "1920x1080, 64 spp, depth 8 per frame", "width": W, "height": H, "spp": spp,
                      "max_depth": depth}}
    ctx = abi.Context(0)
    # flat: one rigid bone per sphere (bone nsph + k: the identity, for the lights)
    h = ctx.upload(s)
    ctx.build_bvh(h)
    nv = s.positions.shape[0]
    ids = np.full((nv, 4), -1, np.int32)
    w = np.zeros((nv, 4), np.float32)
    owner = np.full(nv, nsph, np.int32)
    for i in range(nsph):
        owner[s.indices[first[i]:first[i + 1]].ravel()] = i
    ids[:, 0] = owner
    w[:, 0] = 1.0
    ctx.set_skin(h, ids, w)
    upd, ren = [], []
    for f, xf in enumerate(moves):
        bones = np.concatenate([xf[:nsph], np.eye(4, dtype=np.float32)[None, :3]])
        t1 = time.perf_counter()
        ctx.skin_refit(h, bones)
        t2 = time.perf_counter()
        ctx.render_resident(h, s.camera, W, H, spp, depth, R["rr_depth"], 1, 1, R["seed"])
        t3 = time.perf_counter()
        if f > 0 or frames == 1:
            upd.append((t2 - t1) * 1e3)
            ren.append((t3 - t2) * 1e3)
    flat_img = ctx.film(W * H)
    ctx.release(h)
    out["flat"] = {"update_ms": round(float(np.mean(upd)), 3), "render_ms": round(float(np.mean(ren)), 2),
                   "frame_ms": round(float(np.mean(upd) + np.mean(ren)), 2),
                   "update": "skin (3126 rigid bones over 506k vertices) + refit of the whole BVH"}
    log(f"flat node animation: {out['flat']}")
    # two-level: every sphere its own mesh under one instance
    si = scenes.identity_instanced(s)
    h = ctx.upload(si)
    ctx.build_bvh(h)
    upd, ren = [], []
    for f, xf in enumerate(moves):
        t1 = time.perf_counter()
        ctx.update_instances(h, xf)
        t2 = time.perf_counter()
        ctx.render_resident(h, s.camera, W, H, spp, depth, R["rr_depth"], 1, 1, R["seed"])
        t3 = time.perf_counter()
        if f > 0 or frames == 1:
            upd.append((t2 - t1) * 1e3)
            ren.append((t3 - t2) * 1e3)
    inst_img = ctx.film(W * H)
    ctx.release(h)
    ctx.close()
    rel = np.abs(flat_img[:, :3] - inst_img[:, :3]).mean() / max(float(np.abs(flat_img[:, :3]).mean()), 1e-12)
    out["instanced"] = {"update_ms": round(float(np.mean(upd)), 3), "render_ms": round(float(np.mean(ren)), 2),
                        "frame_ms": round(float(np.mean(upd) + np.mean(ren)), 2),
                        "update": "3129 instance transforms: inverses, instance boxes, top-level refit (host), upload"}
    out["image_mean_rel_diff"] = round(float(rel), 5)
    log(f"instanced node animation: {out['instanced']}")
    print(json.dumps(out), flush=True)


def mlp_mode(args):
    """Radiance MLP against the plain tracer at equal wall time on the Cornell
    box (256x256, depth 5; default) or the 1M-triangle sphere scene (192x108,
    depth 8; --mlp-scene spheres): RMSE against the FP64 oracle's 4096-spp
    image.  The MLP is trained online from the tracer's own vertex-2 samples
    (up to 2^18 per frame), then paths stop at vertex 2 and take its
    prediction."""
    from srt_amd import abi, scenes
    if args.mlp_scene == "spheres":
        gold_path = os.path.join(ROOT, "tests", "golden", "spheres_192x108_fp64_4096spp.npy")
        s = scenes.spheres()
        W, H, depth, rr, train_spp = 192, 108, 8, 4, 64
        workload = ("spheres-1M: 192x108, depth 8, NEE+MIS (RMSE vs FP64 4096-spp oracle); "
                    "MLP 63->4x64->3 ReLU, Adam lr 1e-3")
    else:
        gold_path = os.path.join(ROOT, "tests", "golden", "cornell_256x256_fp64_4096spp.npy")
        s = scenes.cornell()
        W, H, depth, rr, train_spp = 256, 256, 5, 4, 8
        workload = "cornell-36: 256x256, depth 5, NEE+MIS; MLP 63->4x64->3 ReLU, Adam lr 1e-3"
    gold = np.load(gold_path).astype(np.float64)
    ctx = abi.Context(0)
    h = ctx.upload(s)
    ctx.build_bvh(h)

    def rmse(img):
        return float(np.sqrt(np.mean((img[:, :3].astype(np.float64) - gold) ** 2)))

    reps = 4      # independent seeds per comparison point: RMSE = sqrt(mean MSE), low-spp images are firefly-noisy

    def timed_render(spp, seed, mode=0):
        """(root mean MSE over `reps` seeds, the mean image, seconds of one render)"""
        ctx.render(h, s.camera, W, H, spp, depth, rr, 1, 1, seed, mlp_mode=mode)     # warm
        t0 = time.perf_counter()
        img = ctx.render(h, s.camera, W, H, spp, depth, rr, 1, 1, seed + 1, mlp_mode=mode)
        dt = time.perf_counter() - t0
        mse = [rmse(img) ** 2]
        acc = img[:, :3].astype(np.float64)
        for r in range(1, reps):
            im = ctx.render(h, s.camera, W, H, spp, depth, rr, 1, 1, seed + 1 + 7919 * r, mlp_mode=mode)
            mse.append(rmse(im) ** 2)
            acc += im[:, :3]
        return float(np.sqrt(np.mean(mse))), acc / reps, dt

    # online training: frames of up to 2^18 samples harvested from W x H x train_spp tracer paths
    train_frames, steps = args.mlp_frames, args.mlp_steps
    ctx.mlp_create(seed=1, lr=1e-3)
    losses = []
    t0 = time.perf_counter()
    for f in range(train_frames):
        loss, n = ctx.mlp_train_frame(h, s.camera, W, H, train_spp, depth, rr, seed=5000 + f, max_samples=1 << 18,
                                      steps=steps)
        losses.append(loss)
    if args.mlp_decay_frames > 0:        # a few more frames at a tenth of the learning rate
        ctx.mlp_set_hparams(lr=1e-4)
        for f in range(args.mlp_decay_frames):
            loss, n = ctx.mlp_train_frame(h, s.camera, W, H, train_spp, depth, rr, seed=6000 + f,
                                          max_samples=1 << 18, steps=steps)
            losses.append(loss)
    t_train = time.perf_counter() - t0

    # MFMA roofline of the training step: two more training frames (not part of
    # the comparison) with per-kernel HIP-event timing; a frame's n samples are
    # split into `steps` mini-batches, so the steps together see each sample once
    ctx.set_option("profile", 1)
    ctx.stats_reset()
    n_prof = 0
    for f in range(2):
        _, n = ctx.mlp_train_frame(h, s.camera, W, H, train_spp, depth, rr, seed=9000 + f, max_samples=1 << 18,
                                   steps=steps)
        n_prof += n
    ctx.set_option("profile", 0)
    st = ctx.stats()
    train_k = [k for k in ("mlp_train",) if k in st]
    k_ms = sum(st[k]["ms"] for k in train_k)
    flop = MLP_TRAIN_FLOP * n_prof
    achieved = flop / (k_ms * 1e-3) / 1e12 if k_ms > 0 else 0.0
    mlp_roofline = {"bound": "mfma", "kernel": "+".join(train_k), "achieved": round(achieved, 2),
                    "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / MFMA_BF16_PEAK_TFLOPS, 4),
                    "flop_per_sample": MLP_TRAIN_FLOP, "samples": int(n_prof),
                    "ms_per_step": round(k_ms / (2 * steps), 4), "samples_per_step": int(n_prof // (2 * steps)),
                    "adam_ms_per_step": round(st.get("mlp_adam", {}).get("ms", 0.0) / (2 * steps), 4),
                    "per_kernel_ms": {k: round(st[k]["ms"] / (2 * steps), 4) for k in train_k}}
    log(f"MLP training: {achieved:.1f} TFLOP/s over {train_k}, {mlp_roofline['ms_per_step']} ms/step")
    budgets = {}
    for spp in (16, 256):
        r_p, img_p, t_p = timed_render(spp, 100 + spp)                  # plain at this spp
        r_m, img_m, t_m = timed_render(spp, 200 + spp, mode=1)          # MLP form at this spp
        spp_eq = max(1, int(round(spp * t_m / t_p)))                    # plain given the MLP render's time
        r_e, _, t_e = timed_render(spp_eq, 300 + spp)
        spp_tot = max(1, int(round(spp * (t_m + t_train) / t_p)))       # plain given render + all training time
        r_t, _, t_t = timed_render(spp_tot, 400 + spp)
        # training amortised over a sequence of K frames of the scene (one training, K MLP renders)
        amort = {}
        for K in (8, 32, 128):
            spp_k = max(1, int(round(spp * (t_m + t_train / K) / t_p)))
            r_k, _, _ = timed_render(spp_k, 500 + spp + K)
            amort[f"K{K}"] = {"plain_spp": spp_k, "plain_rmse": round(r_k, 5), "mlp_wins": bool(r_m < r_k)}
        budgets[f"{spp}spp"] = {
            "mlp": {"rmse": round(r_m, 5), "ms": round(t_m * 1e3, 2)},
            "plain_same_spp": {"rmse": round(r_p, 5), "ms": round(t_p * 1e3, 2)},
            "plain_equal_render_time": {"spp": spp_eq, "rmse": round(r_e, 5), "ms": round(t_e * 1e3, 2)},
            "plain_equal_time_with_training": {"spp": spp_tot, "rmse": round(r_t, 5), "ms": round(t_t * 1e3, 2)},
            "mlp_wins_equal_render_time": bool(r_m < r_e),
            "mlp_wins_counting_training": bool(r_m < r_t),
            "training_amortised_over_K_frames": amort,
            "mlp_mean_ratio": round(float(img_m.mean() / img_p.mean()), 4),
            "mlp_mean_vs_golden": round(float(img_m.mean() / gold.mean()), 4)}
    # MLP inference rate inside a 256-spp MLP-form render (per-kernel HIP events, query count read back)
    ctx.set_option("profile", 1)
    ctx.stats_reset()
    ctx.render(h, s.camera, W, H, 256, depth, rr, 1, 1, 777, mlp_mode=1)
    ctx.set_option("profile", 0)
    q = ctx.stats().get("mlp_query", {"ms": 0.0, "items": 0})
    q_tf = q["items"] * MLP_FWD_FLOP / (q["ms"] * 1e-3) / 1e12 if q["ms"] > 0 else 0.0
    inference = {"queries": int(q["items"]), "ms": round(q["ms"], 4), "achieved": round(q_tf, 2), "unit": "TFLOP/s",
                 "frac": round(q_tf / MFMA_BF16_PEAK_TFLOPS, 4), "flop_per_query": MLP_FWD_FLOP}
    log(f"MLP inference: {q['items']} queries in {q['ms']:.3f} ms, {q_tf:.1f} TFLOP/s")
    out = {"metric": "RMSE vs FP64 4096-spp oracle at equal wall time (radiance MLP vs plain tracer)",
           "unit": "RMSE", "higher_is_better": False, "dtype": "f32 tracer, bf16 MLP (fp32 accumulation)",
           "data": "synthetic",
           "config": {"workload": workload + f", {train_frames} training frames x {steps} steps, up to 2^18 "
                                             "samples/frame",
                      "golden": os.path.relpath(gold_path, ROOT)},
           "training": {"ms": round(t_train * 1e3, 1), "frames": train_frames, "decay_frames": args.mlp_decay_frames,
                        "loss_first": round(float(losses[0]), 4),
                        "loss_last": round(float(losses[-1]), 4)},
           "roofline": mlp_roofline,
           "inference": inference,
           "budgets": budgets, "rmse_seeds_per_point": reps}
    ctx.close()
    print(json.dumps(out), flush=True)


def serve_load(rate_sat=200.0, window_sat=30.0, window_sub=30.0, seed=20261202, path_pool=0, serve_age=None):
    """Continuous-batching serving under Poisson arrivals over 8 resident
    scenes (3 Cornell variants, 3 1M-triangle sphere scenes with different
    generator seeds, 2 posed skinned tubes); resolution uniform from {256^2,
    512^2, 1280x720, 1920x1080}, spp uniform from {4, 16, 64}; every scene has
    its own radiance MLP, trained online from its own tracer samples, and the
    MLP form is on for each request with probability 1/2.  Run 1: `rate_sat`
    requests/s (200: the offered load of the serving configuration) for
    `window_sat` s, then queued requests are cancelled; reports served
    Mpaths/s.  Run 2: Poisson arrivals at about 70 % of the measured capacity
    for `window_sub` s, drained; reports p50/p99 request latency (submit to
    image in host memory)."""
    from srt_amd import abi, scenes
    rng = np.random.default_rng(seed)
    ctx = abi.Context(0)
    if serve_age is not None:
        ctx.set_option("serve_age", serve_age)
    sc = []
    for fov in (40.0, 35.0, 45.0):
        sc.append(scenes.cornell(fov=fov))
    for sd in (20261202, 20261203, 20261204):
        sc.append(scenes.spheres(seed=sd))
    sc.append(scenes.tube())
    sc.append(scenes.tube(seed=12))
    handles = []
    t0 = time.perf_counter()
    for k, s in enumerate(sc):
        h = ctx.upload(s)
        ctx.build_bvh(h)
        if getattr(s, "skin", None) is not None:
            ctx.set_skin(h, s.skin["bone_ids"], s.skin["bone_weights"])
            ctx.skin_refit(h, s.skin["matrices"](0 if k == 6 else 30))
        handles.append(h)
    t_scenes = time.perf_counter() - t0
    t0 = time.perf_counter()
    for k, (s, h) in enumerate(zip(sc, handles)):
        ctx.mlp_select(h)
        ctx.mlp_create(seed=1 + k, lr=1e-3)
        R = s.render
        for f in range(8):
            ctx.mlp_train_frame(h, s.camera, 256, 256, 8, max(3, R["max_depth"]), R["rr_depth"], seed=7000 + f,
                                max_samples=1 << 18, steps=16)
    ctx.mlp_select(0)
    t_mlp = time.perf_counter() - t0
    log(f"serving: 8 scenes resident in {t_scenes:.2f} s, 8 MLPs trained in {t_mlp:.2f} s")
    res_choices = [(256, 256), (512, 512), (1280, 720), (1920, 1080)]
    spp_choices = [4, 16, 64]

    def draw():
        k = int(rng.integers(len(sc)))
        W, H = res_choices[int(rng.integers(4))]
        spp = spp_choices[int(rng.integers(3))]
        m = 1 if rng.random() < 0.5 else 0
        return k, W, H, spp, m

    def run(rate, window, keep_images, drain):
        ctx.serve_start(path_pool)
        reqs = []
        t0 = time.perf_counter()
        t_next = t0
        offered = 0
        while True:
            t_next += rng.exponential(1.0 / rate)
            if t_next - t0 > window:
                break
            while time.perf_counter() < t_next:
                time.sleep(max(0.0, min(0.002, t_next - time.perf_counter())))
            k, W, H, spp, m = draw()
            s = sc[k]
            R = s.render
            t, _ = ctx.submit(handles[k], s.camera, W, H, spp, R["max_depth"], R["rr_depth"], 1, 1,
                              int(rng.integers(1 << 30)), mlp_mode=m, want_image=keep_images)
            reqs.append((t, W * H * spp))
            offered += W * H * spp
        st_window = ctx.serve_stats()
        t_window = time.perf_counter() - t0
        # drain: every request's result; otherwise those finished within the window
        results = [ctx.wait(t) if drain else ctx.poll(t) for t, _ in reqs]
        ctx.serve_stop(drain=drain)
        ctx.forget_tickets()
        lat = [r["done_ms"] - r["submit_ms"] for r in results if r["status"] == 2]
        return dict(requests=len(reqs), completed=len(lat), offered_Mpaths_s=round(offered / window / 1e6, 1),
                    served_Mpaths_s=round(st_window["served_paths"] / t_window / 1e6, 1), window_s=window,
                    p50_ms=round(float(np.percentile(lat, 50)), 1) if lat else None,
                    p99_ms=round(float(np.percentile(lat, 99)), 1) if lat else None)

    sat = run(rate_sat, window_sat, False, False)
    sat["rate_per_s"] = rate_sat
    log(f"serving at {rate_sat}/s: {sat}")
    mean_paths = np.mean([W * H for W, H in res_choices]) * np.mean(spp_choices)
    sub_rate = max(1.0, 0.7 * sat["served_Mpaths_s"] * 1e6 / mean_paths)
    sub = run(sub_rate, window_sub, True, True)
    sub["rate_per_s"] = round(sub_rate, 1)
    log(f"serving at {sub['rate_per_s']}/s (70 % of capacity): {sub}")
    ctx.close()
    return {"workload": "Poisson arrivals over 8 resident scenes (3 Cornell variants, 3 1M-triangle sphere scenes, "
                        "2 posed 200k-triangle tubes); res U{256^2,512^2,1280x720,1920x1080}, spp U{4,16,64}; "
                        "one radiance MLP per scene, MLP form on p=1/2",
            "saturation": sat, "below_saturation": sub,
            "latency_ms": {"p50": sub["p50_ms"], "p99": sub["p99_ms"], "unit": "ms per request, submit to host image",
                           "at_rate_per_s": sub["rate_per_s"]}}


def serve_mode(args):
    """The serving workload alone, with long windows (default 30 s each)."""
    out = serve_load(args.serve_rate, args.serve_window, args.serve_window, args.serve_seed, serve_age=args.serve_age)
    line = {"metric": "served Mpaths/s at saturation; p50/p99 request latency below saturation",
            "value": out["saturation"]["served_Mpaths_s"], "unit": "Mpaths/s", "n_gpus": 1,
            "higher_is_better": True, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "serving: " + out["workload"], "rate_per_s": args.serve_rate,
                       "window_s": args.serve_window},
            "saturation": out["saturation"], "below_saturation": out["below_saturation"],
            "latency_ms": out["latency_ms"]}
    print(json.dumps(line), flush=True)


def frame_roofline(kern, st, steps, bytes_step, segments):
    """Roofline of the dominant frame kernel against the resource that binds it.

    The traversal kernels are bound by vector-ALU issue, not by HBM: their
    node and triangle fetches mostly hit in L2/MALL.  So `achieved` is the
    kernel's active VALU lane-operations per second (wave64 VALU instructions
    per launch x 64 x active-lane fraction, from the committed SQ counter
    summary, over the live HIP-event launch time) against the published FP32
    vector peak as lane operations (157.3 TFLOP/s = 78.6 T lane-FMA/s: 256 CUs
    x 4 SIMDs x 32 lanes per clock at 2.4 GHz); `frac` = achieved / peak.
    Compares, selects and min / max issue at about half that rate on this GPU,
    and they are about half of a node step's VALU instructions.
    `traffic` is the kernel's calibrated HBM bytes per launch (FETCH_SIZE +
    WRITE_SIZE counters), and `hbm` gives, for each frame kernel, those bytes
    over its live launch time against the 8 TB/s HBM peak.  `algorithmic`
    keeps the byte model (ray and hit records, 112 B per 4-wide node visit,
    36 B per triangle test): it counts cache hits, so its rate is not an HBM
    figure."""
    avg_ms = {k: st[k]["ms"] / st[k]["launches"] for k in bytes_step if k in st and st[k]["launches"]}
    launches = {k: st[k]["launches"] / steps for k in avg_ms}
    hbm = {}
    for k in avg_ms:
        t = measured_traffic(k)
        if t:
            gbs = t / (avg_ms[k] * 1e-3) / 1e9
            hbm[k] = {"bytes_per_launch": t, "avg_launch_ms": round(avg_ms[k], 4), "GBs": round(gbs, 1),
                      "frac": round(gbs / HBM_PEAK_GBS, 4)}
    alg = {k: {"bytes_per_launch": int(bytes_step[k] / launches[k]),
               "GBs": round(bytes_step[k] / launches[k] / (avg_ms[k] * 1e-3) / 1e9, 1)} for k in avg_ms}
    r = {"kernel": kern, "avg_launch_ms": round(avg_ms[kern], 4), "traffic": hbm.get(kern, {}).get("bytes_per_launch"),
         "hbm": hbm, "algorithmic": alg}
    vc = measured_valu_counts(kern)
    if vc:
        insts, lanes, busy, source = vc
        lane_ops = insts * 64.0 * lanes
        achieved = lane_ops / (avg_ms[kern] * 1e-3) / 1e12
        r.update({"bound": "valu", "achieved": round(achieved, 3), "peak": VALU_PEAK_TLANE_OPS,
                  "unit": "Tlane-op/s", "frac": round(achieved / VALU_PEAK_TLANE_OPS, 4),
                  "valu_issue": {"wave_instructions_per_launch": insts, "active_lanes_frac": round(lanes, 3),
                                 "simd_busy_frac": round(busy, 3),
                                 "lane_ops_per_segment": round(lane_ops * launches[kern] / segments, 1),
                                 "source": source}})
    elif kern in hbm:
        r.update({"bound": "hbm", "achieved": hbm[kern]["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": hbm[kern]["frac"]})
    else:
        r.update({"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
                  "note": "no counter summaries under profiles/"})
    return r


def host_cpu():
    """Usable host cores and their description: the scheduler affinity set,
    capped by a cgroup CPU quota when one is set (cgroup v2 cpu.max or v1
    cfs_quota_us), plus os.cpu_count() and the /proc/cpuinfo model name."""
    affinity = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = float(q) / float(per)
    except (OSError, ValueError):
        try:
            q = int(open("/sys/fs/cgroup/cpu/cpu.cfs_quota_us").read())
            per = int(open("/sys/fs/cgroup/cpu/cpu.cfs_period_us").read())
            if q > 0:
                quota = q / per
        except (OSError, ValueError):
            pass
    usable = affinity if quota is None else max(1, min(affinity, int(quota)))
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"usable": usable, "affinity": affinity, "cgroup_quota": quota, "cpu_count": os.cpu_count(),
            "model": model}


def measured_traffic(kernel):
    """Calibrated HBM bytes per launch of `kernel` from the committed counter summary."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as fh:
        for line in fh:
            p = line.split()
            if len(p) == 3 and p[0] == "traffic_bytes_per_launch" and p[1] == kernel:
                return int(p[2])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-row-step", type=int, default=8, help="CPU baseline renders every k-th row")
    ap.add_argument("--path-pool", type=int, default=0)
    ap.add_argument("--mode", choices=["frame", "copy", "skin", "mlp", "serve", "passes", "anim"], default="frame")
    ap.add_argument("--mlp-frames", type=int, default=16)
    ap.add_argument("--mlp-steps", type=int, default=8)
    ap.add_argument("--mlp-decay-frames", type=int, default=0,
                    help="MLP mode: training frames at lr 1e-4 after the first ones")
    ap.add_argument("--mlp-scene", choices=["cornell", "spheres"], default="cornell")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-window", type=float, default=30.0)
    ap.add_argument("--serve-seed", type=int, default=20261202)
    ap.add_argument("--serve-age", type=int, default=None,
                    help="serving order weight of waiting time, per mille (library default when absent)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg of the default line")
    ap.add_argument("--serve-leg-window", type=float, default=5.0,
                    help="seconds of each Poisson window of the default line's serving leg")
    ap.add_argument("--pmc-summary", nargs=4, metavar="DIR")
    ap.add_argument("--kstats-summary", metavar="DIR")
    ap.add_argument("--valu-summary", metavar="DIR")
    ap.add_argument("--kstats-mode", choices=["frame", "passes", "skin", "instanced"], default="frame")
    ap.add_argument("--mlp-summary", nargs=2, metavar="DIR")
    ap.add_argument("--scene", choices=["spheres", "spheres_instanced"], default="spheres",
                    help="frame mode: the flat 1M-triangle field, or the same field as 16 meshes under 3129 "
                         "instances (two-level tree)")
    args = ap.parse_args()
    if args.mlp_summary:
        mlp_summary(*args.mlp_summary)
        return
    if args.pmc_summary:
        pmc_summary(*args.pmc_summary)
        return
    if args.kstats_summary:
        kstats_summary(args.kstats_summary, args.kstats_mode)
        return
    if args.valu_summary:
        valu_summary(args.valu_summary)
        return
    if args.mode == "copy":
        copy_mode()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the frame would shard by "
              "image tiles or spp slices with one gather/all-reduce, not implemented); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    if args.mode != "frame":
        {"skin": skin_mode, "mlp": mlp_mode, "serve": serve_mode, "passes": passes_mode,
         "anim": anim_mode}[args.mode](args)
        return

    from srt_amd import abi, scenes

    log(f"generating the 1M-triangle scene ({args.scene})")
    s = scenes.spheres() if args.scene == "spheres" else scenes.spheres_instanced()
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
    # a second build of the same scene: no first-call costs (allocations, kernel loading)
    t0 = time.perf_counter()
    ctx.build_bvh(h)
    bvh_rebuild_ms = (time.perf_counter() - t0) * 1e3
    info = ctx.bvh_info(h)
    log(f"BVH: {info['num_nodes']} nodes, depth {info['max_depth']}, {info['clusters']} clusters, {bvh_ms:.1f} ms "
        f"(rebuild {bvh_rebuild_ms:.1f} ms)")

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

    # the same frames handed back in host memory (film copied over PCIe each frame)
    host_img = np.zeros((W * H, 4), np.float32)
    n_host = max(1, min(args.steps, 3))
    t1 = time.perf_counter()
    for _ in range(n_host):
        ctx.render(h, s.camera, W, H, spp, depth, R["rr_depth"], 1, 1, R["seed"], out=host_img)
    host_s = (time.perf_counter() - t1) / n_host
    pcie = {"value": round(paths / host_s / 1e6, 2), "unit": "Mpaths/s", "frames": n_host,
            "ms_per_frame": round(host_s * 1e3, 2),
            "note": "srt_render with a host rgba_out: the float32 RGBA film is copied to pageable host memory "
                    "after every frame"}
    log(f"host-output rate {pcie['value']} Mpaths/s ({pcie['ms_per_frame']} ms/frame)")

    # algorithmic bytes per frame per kernel; the first bounce's rays and path
    # state are computed from the slot, so only later segments read a queue entry
    emitted = segments - paths     # queue entries written by shade (= entries read after the first bounce)
    bytes_step = {
        "trace": RAY_IN * emitted + HIT_OUT * segments + NODE * visits["closest_nodes"] + TRI * visits["closest_tris"],
        "shadow": (SHADOW_IN + L_RMW) * shadow_rays + NODE * visits["shadow_nodes"] + TRI * visits["shadow_tris"],
        "shade": (HIT_OUT + TRI_FETCH + MATERIAL) * segments + 2 * QUEUE_ENTRY * emitted + SHADOW_IN * shadow_rays,
    }
    kern = max((k for k in st if k in bytes_step), key=lambda k: st[k]["ms"])
    # the VALU and traffic counters behind the roofline were collected on the flat field
    roofline = frame_roofline(kern, st, args.steps, bytes_step, segments) if args.scene == "spheres" else None

    cpu = None
    if not args.no_cpu:
        from oracle import pyoracle as O   # CPU baseline leg only
        hc = host_cpu()
        threads = hc["usable"]
        log(f"CPU baseline: oracle on every {args.cpu_row_step}th row, {threads} threads on {hc}")
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
               "host": {"usable_cores": hc["usable"], "affinity_cores": hc["affinity"],
                        "cgroup_cpu_quota": hc["cgroup_quota"], "os_cpu_count": hc["cpu_count"],
                        "model": hc["model"]},
               "gpu_vs_cpu": None,
               "sample": f"{len(rows)} of {H} rows (every {args.cpu_row_step}th) of the same 1920x1080x64spp frame, "
                         f"{cpu_paths} paths, {secs:.1f} s",
               "gpu_vs_cpu_max_rel_diff": float(rel.max())}
        cpu["gpu_vs_cpu"] = round(value / cpu["value"], 1)
        log(f"CPU {cpu['value']} Mpaths/s, GPU/CPU {value / cpu['value']:.1f}x, max rel diff {rel.max():.2e}")

    ctx.close()
    serving = None
    if not args.no_serve:
        serving = serve_load(args.serve_rate, args.serve_leg_window, args.serve_leg_window, args.serve_seed,
                             serve_age=args.serve_age)

    lat_sorted = sorted(lat)
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mpaths/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": ("spheres-1M: 3125 frequency-4 icospheres (1,000,000 triangles) + 4 quad lights, "
                                "Disney, 1920x1080, 64 spp, max depth 8, NEE+MIS") if args.scene == "spheres" else
                               ("spheres-1M instanced: 16 frequency-4 icosphere meshes under 3125 instances "
                                "(1,000,000 world triangles) + 4 quad-light instances, Disney, 1920x1080, 64 spp, "
                                "max depth 8, NEE+MIS"),
                   "width": W, "height": H, "spp": spp, "max_depth": depth,
                   "triangles": s.num_triangles if args.scene == "spheres" else 3125 * 320 + 8,
                   "paths_per_step": paths, "parallelism": "single-gpu"},
        "latency_ms": (serving["latency_ms"] if serving else
                       {"p50": round(float(np.percentile(lat_sorted, 50)), 2),
                        "p99": round(float(np.percentile(lat_sorted, 99)), 2), "unit": "ms per frame request"}),
        "frame_latency_ms": {"p50": round(float(np.percentile(lat_sorted, 50)), 2),
                             "p99": round(float(np.percentile(lat_sorted, 99)), 2)},
        "segments_per_path": round(segments / paths, 3), "shadow_rays_per_path": round(shadow_rays / paths, 3),
        "bvh_build_ms": round(bvh_ms, 1),
        "bvh_rebuild_ms": round(bvh_rebuild_ms, 1),
        "pcie_inclusive": pcie,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
