#!/usr/bin/env python3
"""Headline benchmark of the MI355X PM gravity solver.

Workload: 256^3 particles from a Zel'dovich displacement (P(k) ~ k^-2 up to
the mesh Nyquist, 3-D rms displacement 0.5 cell) of a node-aligned lattice,
box L = 256 Mpc/h, 512^3 mesh, equal masses, G = 1, no filter.  One step is
one full force evaluation: bucket sort, fixed-point CIC deposit, R2C FFT,
fused Green's function x gradient, C2R FFTs, CIC interpolation, with the
particle arrays already resident in HBM.  The timed path takes the
positions as the reference's Vec3 f64 array (unrounded; cells formed in FP64
on the device), f64 masses, and writes f64 Vec3 accelerations; the f32 SoA
entry point is timed beside it.

Printed JSON line (stdout, one line):
  value         whole-job M particle-forces/s (N * steps / timed wall time)
  roofline      the slowest stage group's algorithmic bytes (the compulsory
                bytes of that part of the solve, as the survey's measurement
                section defines them) / its mean event-timed duration inside
                the timed region, against 8 TB/s
  cpu_baseline  the FP64 CPU oracle (lean R2C mode, all host cores) on the
                same unrounded f64 inputs
  rms_force_err / max_force_err  vs that FP64 result
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gadget-ng_amd"))

HBM_PEAK_GBPS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stage_bytes(n: int, nm: int, f64: bool = False) -> dict:
    """Bytes each pass of this implementation must move (packed half
    spectrum: one complex field = nm^3/2 complex = 4 nm^3 bytes); with f64
    the particles are Vec3 f64 positions, f64 masses and f64 accelerations."""
    m = nm ** 3
    pin, pos, acc, rec = (32, 24, 24, 24) if f64 else (16, 12, 12, 24)
    return {
        "bucket": pin * n + 8 * n,           # positions, masses in; bucket, rank out
        "sort": pin * n + 8 * n + rec * n,   # positions, masses, bucket, rank in; cell records out
        "deposit": rec * n + 4 * m,          # cell records in; f32 mass mesh out
        "fft_x_fwd": 4 * m + 4 * m,          # mass mesh in, spectrum out
        "fft_y_fwd": 8 * m,                  # spectrum in and out
        "fft_z_green": 4 * m + 8 * m,        # spectrum in, A and B out
        "fft_y_inv": 8 * m + 12 * m,         # A, B in; three fields out
        "fft_x_inv": 12 * m + 12 * m,        # three fields in, three force meshes out
        "interp": pos * n + 12 * m + acc * n,  # positions, force meshes in; accelerations out
    }


# Stage groups of the survey's algorithmic-byte definition and the pipeline
# stages that make up each: deposit chain (bucket count, sort, deposit), R2C
# (x and y forward), k-space (z pass with the Green's function), C2R (y and x
# inverse) and interpolation.
GROUPS = {"deposit_chain": ("bucket", "sort", "deposit"), "r2c": ("fft_x_fwd", "fft_y_fwd"),
          "kspace": ("fft_z_green",), "c2r": ("fft_y_inv", "fft_x_inv"), "interp": ("interp",)}


F32_PARTICLES = (16, 12, 12)   # f32 x, y, z, m in; x, y, z re-read; f32 accelerations out
F64_PARTICLES = (32, 24, 24)   # f64 Vec3 positions and f64 masses; f64 Vec3 accelerations
SOA64_PARTICLES = (28, 24, 12)  # f64 SoA positions and f32 masses; f32 accelerations (the integrator)


def group_bytes(n: int, nm: int, particles=F32_PARTICLES) -> dict:
    """Algorithmic bytes per stage group (M = nm^3, M_h = nm^2 (nm/2 + 1)):
    deposit write 4M, R2C 4M + 8M_h, k-space 8M_h + 24M_h, C2R 24M_h + 12M,
    interpolation read 12M; particles (in, re-read, out) bytes each:
    16, 12, 12 for f32 SoA, so the groups sum to B = 40N + 32M + 64M_h; 32,
    24, 24 for f64 Vec3 positions, f64 masses and f64 accelerations
    (80N + 32M + 64M_h)."""
    m, mh = nm ** 3, nm * nm * (nm // 2 + 1)
    pin, pre, pout = particles
    return {"deposit_chain": pin * n + 4 * m, "r2c": 4 * m + 8 * mh, "kspace": 8 * mh + 24 * mh,
            "c2r": 24 * mh + 12 * m, "interp": 12 * m + pre * n + pout * n}


def merge_chunked(stages: dict, sb: dict, tuning: dict) -> None:
    """With plane chunks a pass pair runs interleaved chunk by chunk and its
    two stages share one timing slot: report the pair as one stage, with the
    sum of both passes' bytes."""
    for on, a, b, name in ((tuning.get("chunk_fwd", 0) and not tuning.get("y_fwd_persistent"), "fft_x_fwd",
                            "fft_y_fwd", "fft_xy_fwd"),
                           (tuning.get("chunk_inv", 0) and not tuning.get("y_inv_persistent"), "fft_y_inv",
                            "fft_x_inv", "fft_yx_inv")):
        if on and a in stages and b in stages:
            stages[name] = stages.pop(a) + stages.pop(b)
            sb[name] = sb.pop(a) + sb.pop(b)


def roofline_of(stages: dict, n: int, nm: int, workload: str, particles=F32_PARTICLES) -> tuple:
    """(roofline object, per-group table) of one solve's stage times."""
    gb = group_bytes(n, nm, particles)
    gt = group_times(stages)
    dom = max(gt, key=gt.get)
    achieved = gb[dom] / (gt[dom] * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic(dom, nm, workload)
    roof = {"bound": "hbm", "kernel": dom, "kernels": list(GROUP_KERNELS[dom]), "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": traffic, "traffic_source": traffic_src, "algorithmic_bytes": gb[dom],
            "mean_ms": round(gt[dom], 5)}
    groups = {g: {"ms": round(gt[g], 5), "algorithmic_bytes": gb[g],
                  "frac": round(gb[g] / (gt[g] * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4) if gt[g] > 0 else None}
              for g in gb}
    return roof, groups


def group_times(stages: dict) -> dict:
    return {g: sum(stages.get(s, 0.0) for s in members) for g, members in GROUPS.items()}


# kernel-name prefixes of each stage group's launches (the k-space stage also
# launches the small k_x = 0 column kernel, counted with it)
GROUP_KERNELS = {"deposit_chain": ("k_bucket_count", "k_scan_local", "k_scan_partials", "k_scan_add", "k_scatter",
                                   "k_deposit_tiles"),
                 "r2c": ("k_fft_x_fwd", "k_fft_y_fwd"), "kspace": ("k_fft_z_green",),
                 "c2r": ("k_fft_y_inv", "k_fft_x_inv"), "interp": ("k_interp",)}


def measured_traffic(group: str, nm: int, workload: str):
    """HBM bytes per solve of a stage group's kernels (FETCH_SIZE +
    WRITE_SIZE, each scaled by what the counter reads for a 1 GiB copy) from
    the committed rocprofv3 summary of the same workload, with the file and
    the commit it was measured at; (None, None) when there is none."""
    name = {"zeldovich": "pm_headline_256cubed_512mesh_traffic.json",
            "clustered": "pm_clustered_512cubed_1024mesh_traffic.json"}.get(workload)
    path = os.path.join(ROOT, "profiles", name) if name else None
    if not path or not os.path.exists(path):
        return None, None
    try:
        doc = json.load(open(path))
        kern = doc["kernels"]
    except (OSError, ValueError, KeyError):
        return None, None
    total, found = 0.0, False
    for k, v in kern.items():
        short = k.split("::")[-1]
        if short.startswith(GROUP_KERNELS[group]) and (f"<{nm}" in short or "<" not in short):
            # bytes per solve: a chunked pass runs several launches per solve
            total += v.get("per_solve_bytes", v["read_bytes"] + v["write_bytes"])
            found = True
    src = f"profiles/{name}" + (f" (measured at commit {doc['commit']})" if doc.get("commit") else "")
    return (int(total) if found else None), src


def profile_summary(args) -> int:
    """Per-kernel table of a rocprofv3 run: calls, mean/min/max duration from
    the kernel statistics, HBM bytes per dispatch from FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch), each scaled by what the counter reads for
    the largest device-to-device copy in the same run (the 1 GiB copy of
    --copy-calibration), and per-solve bytes for chunked passes."""
    import collections
    import csv
    import glob

    kt_dir, fetch_dir, write_dir = args.profile_summary

    def one(pattern, d):
        found = sorted(glob.glob(os.path.join(d, "**", pattern), recursive=True))
        if not found:
            raise SystemExit(f"no {pattern} under {d}")
        return found[0]

    def short(name):
        name = name.replace("void ", "")
        i = name.find("(")
        return name[:i] if i > 0 else name

    stats = {r["Name"]: r for r in csv.DictReader(open(one("*kernel_stats.csv", kt_dir)))}

    def counter(d, which):
        rows = [r for r in csv.DictReader(open(one("*counter_collection.csv", d))) if r["Counter_Name"] == which]
        per = collections.defaultdict(list)
        for r in rows:
            per[short(r["Kernel_Name"])].append(float(r["Counter_Value"]))
        copy = max(float(r["Counter_Value"]) for r in rows if "copyBuffer" in r["Kernel_Name"])
        scale = float(1 << 30) / (copy * 1024.0)
        return {k: sum(v) / len(v) * 1024.0 * scale for k, v in per.items() if "copyBuffer" not in k}, scale

    reads, ff = counter(fetch_dir, "FETCH_SIZE")
    writes, wf = counter(write_dir, "WRITE_SIZE")
    solves = sum(int(r["Calls"]) for n, r in stats.items()
                 if short(n).split("::")[-1].startswith(args.summary_solve_kernel) and "col0" not in n)
    rows, kernels = [], {}
    for n, r in stats.items():
        sn = short(n)
        calls = int(r["Calls"])
        mean = float(r["AverageNs"]) / 1e3
        per_solve = calls / solves if solves else 1.0
        rb, wb = reads.get(sn), writes.get(sn)
        rows.append((float(r["TotalDurationNs"]), sn, calls, per_solve, mean, float(r["MinNs"]) / 1e3,
                     float(r["MaxNs"]) / 1e3, rb, wb))
        if rb is not None and wb is not None:
            kernels[sn] = {"mean_us": mean, "read_bytes": rb, "write_bytes": wb, "calls": calls,
                           "per_solve_bytes": (rb + wb) * per_solve, "per_solve_us": mean * per_solve}
    rows.sort(reverse=True)
    lines = [f"# {args.summary_title}", "", "Command (from the repository root, on the GPU box):", "", "```",
             args.summary_command, "```", ""]
    if args.summary_commit:
        lines += [f"Measured at commit {args.summary_commit}.", ""]
    lines += [f"HBM traffic scaling from a 1 GiB device-to-device copy in the same run: FETCH_SIZE reads "
              f"{1 / ff:.3f} of the true bytes (factor {ff:.3f} applied), WRITE_SIZE reads {1 / wf:.3f} "
              f"(factor {wf:.3f}).", "",
              "FETCH_SIZE and WRITE_SIZE count traffic between L2 and the memory fabric, so bytes served by the "
              "256 MiB Infinity Cache count as HBM bytes here.  A chunked pass runs several launches per solve "
              "(column 'per solve'; solves counted by the once-per-solve z pass).", "",
              "| kernel | calls | per solve | mean us | min us | max us | HBM read MiB | HBM write MiB | TB/s |",
              "|---|---|---|---|---|---|---|---|---|"]
    for _, sn, calls, ps, mean, mn, mx, rb, wb in rows:
        tbs = ((rb or 0.0) + (wb or 0.0)) / (mean * 1e-6) / 1e12 if rb is not None else float("nan")
        lines.append(f"| `{sn}` | {calls} | {ps:.2f} | {mean:.1f} | {mn:.1f} | {mx:.1f} | "
                     f"{(rb or 0.0) / 2 ** 20:.1f} | {(wb or 0.0) / 2 ** 20:.1f} | {tbs:.2f} |")
    if args.summary_md:
        with open(args.summary_md, "w") as f:
            f.write("\n".join(lines) + "\n")
    if args.summary_json:
        with open(args.summary_json, "w") as f:
            json.dump({"command": args.summary_command, "commit": args.summary_commit, "fetch_factor": ff,
                       "write_factor": wf, "solves": solves, "kernels": kernels}, f, indent=1)
    print("\n".join(lines))
    return 0


def main() -> int:
    import faulthandler

    faulthandler.enable()  # a crash prints the Python stack of the call it happened in
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=("zeldovich", "clustered", "ensemble", "learned", "kdk", "pk", "ic"),
                    default="zeldovich",
                    help="zeldovich: 256^3 lattice on a 512^3 mesh (the headline); clustered: 512^3 particles, "
                         "70%% lattice + 30%% in 2000 Plummer spheres, 1024^3 mesh, TreePM long range r_s = 1.25 cells")
    ap.add_argument("--n-side", type=int, default=256)
    ap.add_argument("--mesh", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true", help="skip the FP64 CPU oracle (and the error check)")
    ap.add_argument("--cpu-runs", type=int, default=5, help="timed CPU oracle runs after one warm-up")
    ap.add_argument("--no-cpu-faithful", action="store_true",
                    help="skip timing the reference-Rayon-faithful CPU mode (per-thread deposit grids, "
                         "single-threaded Z2Z passes; about 10 s at 512^3, 18 GB of host memory)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch choices")
    ap.add_argument("--no-fp64-form", action="store_true",
                    help="skip timing the FP64 GPU form of the reference algorithm")
    ap.add_argument("--train-boxes", type=int, default=200, help="learned: training seeds 1..N (64^3 boxes)")
    ap.add_argument("--train-steps", type=int, default=2000, help="learned: Adam steps")
    ap.add_argument("--eval-boxes", type=int, default=20, help="learned: held-out 256^3 seeds from 1001")
    ap.add_argument("--only-headline", action="store_true",
                    help="time the f64 entry point only (no f32 entry, host entry points or copy rate): for "
                         "profiler runs, whose per-solve figures need one kind of solve")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 1 GiB device-to-device copy (a kernel of known traffic for scaling "
                         "profiler byte counters)")
    ap.add_argument("--profile-summary", nargs=3, metavar=("KT_DIR", "FETCH_DIR", "WRITE_DIR"),
                    help="no GPU work: summarise three rocprofv3 output folders (--kernel-trace --stats, "
                         "--pmc FETCH_SIZE, --pmc WRITE_SIZE) into --summary-md and --summary-json")
    ap.add_argument("--summary-md", help="profile summary: markdown table to write")
    ap.add_argument("--summary-json", help="profile summary: per-kernel traffic (read by roofline.traffic)")
    ap.add_argument("--summary-title", default="rocprofv3 summary")
    ap.add_argument("--summary-command", default="", help="the profiled command, quoted in the summary")
    ap.add_argument("--summary-commit", default="", help="the commit the profile was measured at")
    ap.add_argument("--summary-solve-kernel", default="k_fft_z_green",
                    help="kernel launched once per solve (counts the solves)")
    args = ap.parse_args()

    if args.profile_summary:
        return profile_summary(args)

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this PM solver has no multi-GPU path "
              "(the slab decomposition is a design, not built); run with --gpus 1", file=sys.stderr)
        return 2
    if not 1 <= args.steps <= 64:
        print("--steps must be in [1, 64] (the stage-timing ring holds 64 solves)", file=sys.stderr)
        return 2

    if args.workload == "learned":
        return bench_learned(args)
    if args.workload == "ensemble":
        return bench_ensemble(args)
    if args.workload == "kdk":
        return bench_kdk(args)
    if args.workload == "pk":
        return bench_pk(args)
    if args.workload == "ic":
        return bench_ic(args)

    import numpy as np

    import gadget_ng_pm_hip as pm
    from gadget_ng_pm_hip import workloads

    t0 = time.perf_counter()
    r_split = 0.0
    if args.workload == "zeldovich":
        nm, n_side = args.mesh, args.n_side
        box = float(n_side)  # 1 Mpc/h lattice spacing
        parts = workloads.zeldovich_lattice(n_side=n_side, nm=nm, box=box, spectral_index=-2.0, rms_disp_cells=0.5,
                                            seed=workloads.SEED_ZELDOVICH, workers=-1, keep_f64=True)
        wl_desc = (f"Zel'dovich lattice {n_side}^3 particles on {nm}^3 mesh, L={box:g} Mpc/h, P(k)~k^-2, "
                   f"rms displacement 0.5 cell, seed 0x6E670002, G=1, no filter; f64 Vec3 positions")
    else:
        nm, box = 1024, 512.0
        parts = workloads.clustered_box(workers=-1, keep_f64=True)
        r_split = 1.25 * box / nm  # reference convention exp(-k^2 r_s^2 / 2)
        wl_desc = ("512^3 particles: 454^3 Zel'dovich lattice + 40641064 in 2000 Plummer spheres (a=0.2 Mpc/h), "
                   f"L=512 Mpc/h, 1024^3 mesh, TreePM long range r_s=1.25 cells, seed 0x6E670003, G=1; "
                   f"f64 Vec3 positions")
    n = parts.n
    pos64 = np.ascontiguousarray(parts.p64)
    mass64 = parts.mass64()
    log(f"[bench] workload generated in {time.perf_counter() - t0:.1f}s: {parts.name}, N={n}")

    # inputs resident in HBM before the timed region (library-owned device buffers)
    d_pos = pm.DeviceArray.from_host(pos64)
    d_mass = pm.DeviceArray.from_host(mass64)
    d_out = pm.DeviceArray((n, 3), np.float64)
    plan = pm.Plan(nm, box, r_split=r_split, device=0)
    t1 = time.perf_counter()
    # the autotuner times whole solves of the workload's own particles
    tuning = plan.autotune_f64(d_pos.ptr, d_mass.ptr, n, d_out.ptr, reps=5) if not args.no_autotune \
        else plan.tuning()
    log(f"[bench] launch choices ({time.perf_counter() - t1:.2f} s): {tuning}")

    if args.copy_calibration:
        src = pm.DeviceArray((1 << 28,), np.float32)
        dst = pm.DeviceArray((1 << 28,), np.float32)
        pm.device_copy(dst, src)
        pm.device_synchronize()
        del src, dst

    def timed(step, steps, warmup):
        for _ in range(warmup):
            step()
        pm.device_synchronize()
        plan.set_timing(True)
        pm.device_synchronize()
        t_start = time.perf_counter()
        for _ in range(steps):
            step()
        pm.device_synchronize()
        wall = time.perf_counter() - t_start
        st = plan.stage_times()
        plan.set_timing(False)
        return wall, st

    # the headline: f64 Vec3 positions and masses in, f64 Vec3 accelerations out
    wall, stages = timed(lambda: plan.solve_device_f64(d_pos.ptr, d_mass.ptr, n, d_out.ptr, 1.0), args.steps,
                         args.warmup)
    gpu_ms = stages.pop("total")
    ms_per_step = wall * 1e3 / args.steps
    value = n * args.steps / wall / 1e6
    log(f"[bench] {args.steps} steps (f64 Vec3 in/out): {ms_per_step:.3f} ms/step wall, {gpu_ms:.3f} ms/step GPU "
        f"events, {value:.1f} M particle-forces/s")
    log("[bench] stage ms: " + ", ".join(f"{k}={v:.4f}" for k, v in stages.items()))
    acc = d_out.numpy()

    # every stage group against its algorithmic bytes, and each implementation
    # pass against the bytes that pass itself must move
    roofline, groups = roofline_of(stages, n, nm, args.workload, F64_PARTICLES)
    achieved = roofline["achieved"]
    sb = stage_bytes(n, nm, f64=True)
    merge_chunked(stages, sb, tuning)
    per_pass = {k: {"ms": round(v, 5), "bytes": sb[k], "gbps": round(sb[k] / (v * 1e-3) / 1e9, 1) if v > 0 else None}
                for k, v in stages.items()}
    b_alg = sum(g["algorithmic_bytes"] for g in groups.values())
    pipeline_gbps = b_alg / (gpu_ms * 1e-3) / 1e9

    if args.only_headline:
        print(json.dumps({"metric": "profiler run", "value": round(value, 2), "unit": "M particle-forces/s",
                          "ms_per_step": round(ms_per_step, 4), "roofline": roofline}), flush=True)
        plan.close()
        return 0

    # the f32 SoA entry point (the hip_pm_solve data layout) on the rounded positions
    x, y, z, m = (pm.DeviceArray.from_host(a) for a in (parts.x, parts.y, parts.z, parts.m))
    ax, ay, az = (pm.DeviceArray((n,), np.float32) for _ in range(3))
    wall32, stages32 = timed(lambda: plan.solve_device(x.ptr, y.ptr, z.ptr, m.ptr, n, ax.ptr, ay.ptr, az.ptr, 1.0),
                             args.steps, args.warmup)
    ms32 = wall32 * 1e3 / args.steps
    acc32 = np.stack([ax.numpy(), ay.numpy(), az.numpy()], axis=1).astype(np.float64)
    for a in (x, y, z, m, ax, ay, az):
        a.free()
    log(f"[bench] f32 SoA entry point: {ms32:.3f} ms/step")

    # drop-in rates through the host C ABI (H2D + D2H included), writing into
    # output arrays the caller keeps across calls (the reference's solvers
    # fill the caller's `out: &mut [Vec3]`); a fresh output array per call
    # adds the page faults of its first touch
    t_e2e, t_e2e32, t_fresh = [], [], []
    host_out = np.empty((n, 3), dtype=np.float64)
    out32 = np.empty((3, n), dtype=np.float32)
    plan.accelerations_f64(pos64, mass64, 1.0, out=host_out)
    for _ in range(3):
        t1 = time.perf_counter()
        plan.accelerations_f64(pos64, mass64, 1.0, out=host_out)
        t_e2e.append(time.perf_counter() - t1)
        t1 = time.perf_counter()
        plan.accelerations_f32(parts.x, parts.y, parts.z, parts.m, 1.0, out=out32)
        t_e2e32.append(time.perf_counter() - t1)
        t1 = time.perf_counter()
        plan.accelerations_f64(pos64, mass64, 1.0)
        t_fresh.append(time.perf_counter() - t1)
    e2e_ms = statistics.median(t_e2e) * 1e3
    e2e32_ms = statistics.median(t_e2e32) * 1e3
    fresh_ms = statistics.median(t_fresh) * 1e3

    result = {
        "metric": "M particle-forces/s + HBM GB/s at 256^3 part/512^3 mesh; rms force err vs FP64"
                  if args.workload == "zeldovich" else "M particle-forces/s, TreePM long range, 512^3 clustered "
                                                       "particles on a 1024^3 mesh; rms force err vs FP64",
        "value": round(value, 2),
        "unit": "M particle-forces/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wl_desc, "particles": n, "mesh": nm, "box": box, "r_split": r_split,
                   "precision": "f64 positions -> FP64 cell index and 43-bit fraction; FP64 weights, int64 "
                                "deposit; FP32 mesh/FFT/interp", "launch_choices": tuning},
        "roofline": roofline,
        "stage_groups": groups,
        "per_pass": per_pass,
        "hbm_gbps": round(pipeline_gbps, 1),
        "hbm_frac": round(pipeline_gbps / HBM_PEAK_GBPS, 4),
        "algorithmic_bytes": b_alg,
        "gpu_ms_per_step": round(gpu_ms, 4),
        "f32_entry": {"ms_per_step": round(ms32, 4), "value": round(n / (ms32 * 1e-3) / 1e6, 2),
                      "gpu_ms_per_step": round(stages32["total"], 4)},
        "pcie_inclusive_ms": round(e2e_ms, 3),
        "pcie_inclusive_ms_f32_entry": round(e2e32_ms, 3),
        "pcie_inclusive_ms_fresh_out": round(fresh_ms, 3),
        "cpu_baseline": None,
        "rms_force_err": None,
        "max_force_err": None,
    }

    # a plain 1 GiB device-to-device copy: the practical HBM ceiling on this box
    cbytes = 1 << 30
    ca, cb = pm.DeviceArray((cbytes,), np.uint8), pm.DeviceArray((cbytes,), np.uint8)
    pm.device_copy(cb, ca)
    pm.device_synchronize()
    t_copy = []
    for _ in range(5):
        t1 = time.perf_counter()
        pm.device_copy(cb, ca)
        pm.device_synchronize()  # device-to-device copies return before they finish
        t_copy.append(time.perf_counter() - t1)
    ca.free()
    cb.free()
    copy_gbps = 2 * cbytes / min(t_copy) / 1e9
    result["copy_gbps"] = round(copy_gbps, 1)
    result["hbm_frac_of_copy"] = round(pipeline_gbps / copy_gbps, 4)
    result["roofline"]["frac_of_copy"] = round(achieved / copy_gbps, 4)

    if not args.no_fp64_form and nm & (nm - 1) == 0:
        # the reference algorithm in FP64 on the same GPU: device-resident
        # (kernels only) and through its host entry point (f64 copies in and
        # out on every call, as the reference's GPU crates do)
        plan.close()
        for a in (d_pos, d_mass, d_out):
            a.free()
        p64 = pm.Plan64(nm, box, r_split=r_split, device=0)
        e_pos = pm.DeviceArray.from_host(pos64)
        e_mass = pm.DeviceArray.from_host(mass64)
        e_out = pm.DeviceArray((n, 3), np.float64)
        p64.solve_device(e_pos.ptr, e_mass.ptr, n, e_out.ptr, 1.0)
        pm.device_synchronize()
        t_dev = []
        for _ in range(3):
            t1 = time.perf_counter()
            p64.solve_device(e_pos.ptr, e_mass.ptr, n, e_out.ptr, 1.0)
            pm.device_synchronize()
            t_dev.append(time.perf_counter() - t1)
        a64 = e_out.numpy()
        for a in (e_pos, e_mass, e_out):
            a.free()
        t_host = []
        for _ in range(2):
            t1 = time.perf_counter()
            p64.accelerations(pos64, mass64, 1.0)
            t_host.append(time.perf_counter() - t1)
        p64.close()
        ms_dev = statistics.median(t_dev) * 1e3
        ms_host = statistics.median(t_host) * 1e3
        result["fp64_gpu_form"] = {
            "ms_device": round(ms_dev, 3), "ms_host_round_trip": round(ms_host, 3),
            "fp32_speedup_device": round(ms_dev / ms_per_step, 2),
            "fp32_speedup_host_vs_host": round(ms_host / e2e_ms, 2),
            "fp32_speedup_vs_host_round_trip": round(ms_host / ms_per_step, 2),
            "fp32_rms_vs_fp64_form": float(f"{np.sqrt(((acc - a64) ** 2).sum() / (a64 ** 2).sum()):.4e}")}
        log(f"[bench] FP64 GPU form: {ms_dev:.2f} ms device-resident, {ms_host:.2f} ms with host copies")

    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle  # CPU baseline and checker only

        oracle.build()
        cores = oracle.get_threads()
        t1 = time.perf_counter()
        ref = oracle.pm_accelerations(pos64, mass64, box, nm, 1.0, r_split=r_split, mode=oracle.MODE_LEAN)
        warm = time.perf_counter() - t1
        times = []
        for _ in range(args.cpu_runs):
            t1 = time.perf_counter()
            oracle.pm_accelerations(pos64, mass64, box, nm, 1.0, r_split=r_split, mode=oracle.MODE_LEAN)
            times.append(time.perf_counter() - t1)
        tc = statistics.median(times) if times else warm
        result["cpu_baseline"] = {
            "value": round(n / tc / 1e6, 3), "unit": "M particle-forces/s", "cores": cores, "kind": "port",
            "sample": f"the full workload ({n} particles, {nm}^3 mesh, unrounded f64 positions), FP64 oracle "
                      f"lean R2C mode, median of {len(times)} timed runs after 1 warm-up ({tc:.2f} s each)"}

        def errs(a, r):
            d = a - r
            nr = np.linalg.norm(r, axis=1)
            floor = 0.05 * nr.mean()
            rel = np.linalg.norm(d, axis=1) / np.maximum(np.maximum(nr, floor), 1e-300)
            return float(f"{np.sqrt((d ** 2).sum() / (r ** 2).sum()):.4e}"), rel
        result["rms_force_err"], rel = errs(acc, ref)
        result["max_force_err"] = float(f"{rel.max():.4e}")
        result["p99_force_err"] = float(f"{np.percentile(rel, 99):.4e}")
        result["rms_force_err_host_path"] = errs(host_out, ref)[0]
        # the f32 entry point: against the oracle on its own (rounded) inputs
        # and against the unrounded f64 truth (the rounding's own share)
        ref32 = oracle.pm_accelerations(parts.pos64(), mass64, box, nm, 1.0, r_split=r_split,
                                        mode=oracle.MODE_LEAN)
        result["f32_entry"]["rms_vs_rounded_inputs"] = errs(acc32, ref32)[0]
        result["f32_entry"]["rms_vs_unrounded_inputs"] = errs(acc32, ref)[0]
        result["f32_entry"]["rounding_alone_rms"] = errs(ref32, ref)[0]
        log(f"[bench] CPU oracle {tc:.2f}s on {cores} threads; rms err {result['rms_force_err']:.3e}, "
            f"max {result['max_force_err']:.3e}; f32 entry rms {result['f32_entry']['rms_vs_rounded_inputs']:.3e} "
            f"vs rounded, {result['f32_entry']['rms_vs_unrounded_inputs']:.3e} vs unrounded inputs")
        if not args.no_cpu_faithful and args.workload == "zeldovich":
            # the reference-Rayon-faithful mode: per-thread deposit grids and
            # single-threaded Z2Z passes (cic.rs:355-401, fft_poisson.rs:356-373)
            ft = []
            for _ in range(args.cpu_runs):
                t1 = time.perf_counter()
                oracle.pm_accelerations(pos64, mass64, box, nm, 1.0, r_split=r_split, mode=oracle.MODE_FAITHFUL,
                                        deposit_mode=oracle.DEPOSIT_PRIVATE_GRIDS)
                ft.append(time.perf_counter() - t1)
            tf = statistics.median(ft)
            result["cpu_faithful"] = {"value": round(n / tf / 1e6, 4), "unit": "M particle-forces/s",
                                      "cores": cores, "seconds": round(tf, 2), "runs": len(ft),
                                      "sample": "the full workload, FP64 oracle in the reference-Rayon-faithful "
                                                "mode (per-thread deposit grids, single-threaded Z2Z passes), "
                                                "median of the timed runs"}
            log(f"[bench] CPU oracle faithful mode: {tf:.1f} s (median of {len(ft)})")
        # calibration against the reference's only published PM timings (AP-08 input,
        # serial CPU PM: 1.83 ms at grid 32, 22.4 ms at grid 64), with this GPU beside it
        ap = {}
        threads = oracle.get_threads()
        oracle.set_threads(1)
        for grid, published in ((32, 1.83), (64, 22.4)):
            q = workloads.ap08_curve(grid)
            p64, m64 = q.pos64(), q.mass64()
            oracle.pm_accelerations(p64, m64, q.box, grid, 1.0, mode=oracle.MODE_FAITHFUL,
                                    deposit_mode=oracle.DEPOSIT_SERIAL)
            tt = []
            for _ in range(5):
                t1 = time.perf_counter()
                oracle.pm_accelerations(p64, m64, q.box, grid, 1.0, mode=oracle.MODE_FAITHFUL,
                                        deposit_mode=oracle.DEPOSIT_SERIAL)
                tt.append(time.perf_counter() - t1)
            gp = pm.Plan(grid, q.box, device=0)
            gp.accelerations_f64(p64, m64, 1.0)
            tg = []
            for _ in range(20):
                t1 = time.perf_counter()
                gp.accelerations_f64(p64, m64, 1.0)
                tg.append(time.perf_counter() - t1)
            gp.close()
            ap[f"grid{grid}"] = {"n": q.n, "reference_published_cpu_ms": published,
                                 "oracle_serial_faithful_ms": round(statistics.median(tt) * 1e3, 3),
                                 "gpu_host_entry_ms": round(statistics.median(tg) * 1e3, 4)}
        oracle.set_threads(threads)
        result["ap08_calibration"] = ap

    print(json.dumps(result), flush=True)
    plan.close()
    return 0


def bench_ensemble(args) -> int:
    """Ensemble workload: 64 independent boxes of uneven size solved together
    through one captured hipGraph; compared with the same boxes solved one
    after another on one stream, and checked per box against the FP64 oracle."""
    import numpy as np

    import gadget_ng_pm_hip as pm
    from gadget_ng_pm_hip import workloads

    boxes, off, x, y, z, m = workloads.ensemble(workers=-1)
    n = int(off[-1])
    d = [pm.DeviceArray.from_host(a) for a in (x, y, z, m)]
    out = [pm.DeviceArray((n,), np.float32) for _ in range(3)]
    batch = pm.Batch(boxes, device=0)
    t0 = time.perf_counter()
    batch.autotune(reps=3)
    t_tune = time.perf_counter() - t0
    tuned = {}
    for i, b in enumerate(boxes):
        tuned[str(b["grid_size"])] = batch.tuning(i)
    log(f"[bench] autotuned {len(tuned)} mesh sizes in {t_tune:.1f} s: {tuned}")

    def graph_step():
        batch.solve_device(d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, off, 1.0, out[0].ptr, out[1].ptr, out[2].ptr)

    def time_steps(steps):
        for _ in range(args.warmup):
            graph_step()
        pm.device_synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            graph_step()
        pm.device_synchronize()
        return time.perf_counter() - t0

    # launch split: every stage one launch over up to k boxes of a mesh side
    # (k = 2^30: one launch per side; 0: one pipeline per box), each timed,
    # then the library's own tuner picks on the same particles
    per_side = {}
    for b in boxes:
        per_side[b["grid_size"]] = per_side.get(b["grid_size"], 0) + 1
    biggest = max(per_side.values())
    split_ms = {}
    for k in sorted({1 << 30, (biggest + 1) // 2, (biggest + 3) // 4, 0}, reverse=True):
        batch.set_box_tables(k)
        split_ms["one launch per side" if k == 1 << 30 else ("per box" if k == 0 else f"{k} boxes per launch")] = \
            round(time_steps(max(args.steps // 2, 3)) * 1e3 / max(args.steps // 2, 3), 4)
    t0 = time.perf_counter()
    chosen = batch.tune_launches(d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, off, 1.0, out[0].ptr, out[1].ptr,
                                 out[2].ptr, reps=3)
    log(f"[bench] launch split ms per ensemble: {split_ms}; tuner chose {chosen} "
        f"({time.perf_counter() - t0:.1f} s)")
    wall = time_steps(args.steps)
    acc = np.stack([o.numpy() for o in out], axis=1).astype(np.float64)

    # the same boxes, one plan each, launched one after another without a graph
    plans = [pm.Plan(b["grid_size"], b["box_size"], device=0) for b in boxes]
    fl = 4

    def serial_step():
        for i, p in enumerate(plans):
            o0 = int(off[i]) * fl
            p.solve_device(d[0].ptr + o0, d[1].ptr + o0, d[2].ptr + o0, d[3].ptr + o0, int(off[i + 1] - off[i]),
                           out[0].ptr + o0, out[1].ptr + o0, out[2].ptr + o0, 1.0)

    serial_step()
    pm.device_synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        serial_step()
    pm.device_synchronize()
    wall_serial = time.perf_counter() - t0

    worst = None
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle  # checker and CPU baseline only

        oracle.build()
        worst = 0.0
        t_cpu = 0.0
        for i, b in enumerate(boxes):
            p = b["particles"]
            t1 = time.perf_counter()
            ref = oracle.pm_accelerations(p.pos64(), p.mass64(), b["box_size"], b["grid_size"], 1.0)
            t_cpu += time.perf_counter() - t1
            got = acc[off[i]:off[i + 1]]
            worst = max(worst, float(np.sqrt(((got - ref) ** 2).sum() / (ref ** 2).sum())))
        cpu = {"value": round(n / t_cpu / 1e6, 3), "unit": "M particle-forces/s", "cores": oracle.get_threads(),
               "kind": "port", "sample": f"every box of the ensemble once ({n} particles), FP64 oracle lean R2C "
                                         f"mode, {t_cpu:.2f} s in total"}
    ms = wall * 1e3 / args.steps
    b_alg = sum(workloads.algorithmic_bytes(int(off[i + 1] - off[i]), b["grid_size"]) for i, b in enumerate(boxes))
    ach = b_alg / (ms * 1e-3) / 1e9
    sizes = {}
    for b in boxes:
        sizes[b["grid_size"]] = sizes.get(b["grid_size"], 0) + 1
    result = {
        "metric": "M particle-forces/s, ensemble of 64 uneven periodic boxes (32^3..128^3 particles) in one hipGraph",
        "value": round(n * args.steps / wall / 1e6, 2), "unit": "M particle-forces/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "64 boxes, N_side from {32,48,64,96,128}, mesh = 2 N_side, L = N_side Mpc/h, "
                               "Zel'dovich P(k)~k^n, n in [-2.5,-1.5], seeds 0x6E670004 + box",
                   "particles": n, "meshes": {str(k): v for k, v in sorted(sizes.items())}},
        "roofline": {"bound": "hbm", "kernel": "whole ensemble (every stage of every box)", "achieved": round(ach, 1),
                     "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBPS, 4), "traffic": None,
                     "algorithmic_bytes": b_alg, "mean_ms": round(ms, 4)},
        "cpu_baseline": cpu,
        "hbm_gbps": round(ach, 1),
        "autotune": {"seconds": round(t_tune, 2), "choices": tuned},
        "launch_split_ms": split_ms,
        "launch_split_chosen": chosen,
        "serial_ms_per_step": round(wall_serial * 1e3 / args.steps, 4),
        "graph_speedup": round(wall_serial / wall, 3),
        "worst_box_rms_force_err": worst,
    }
    print(json.dumps(result), flush=True)
    return 0


def bench_kdk(args) -> int:
    """Device-resident KDK leapfrog: the headline particles (256^3 Zel'dovich
    lattice on a 512^3 mesh) with growing-mode momenta, LambdaCDM background,
    K steps of 2 PM forces each with the state in HBM.  Beside it: the same
    step driven from the host as the reference's GPU path does (positions to
    f32 and over PCIe, accelerations back, f64 kick/drift on the host), the
    FP64 CPU oracle on one step, and both forms at the AP-08 mesh sizes."""
    import numpy as np

    import gadget_ng_pm_hip as pm
    from gadget_ng_pm_hip import integrate as gi
    from gadget_ng_pm_hip import workloads

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import integrate_oracle as io  # the host-driven step and the CPU baseline only

    nm, n_side = args.mesh, args.n_side
    box = float(n_side)
    om, ol = 0.3, 0.7
    h0 = (8.0 * 3.141592653589793 / (3.0 * om)) ** 0.5  # G = 1, mean density 1
    cosmo = gi.CosmologyParams(om, ol, h0)
    c_tuple = (om, ol, h0, -1.0, 0.0, 0.0)
    a0 = 0.02
    hubble = h0 * (om / a0 ** 3 + ol) ** 0.5
    dt = 0.01 / hubble  # about 1 % growth of a per step
    parts = workloads.zeldovich_lattice(n_side=n_side, nm=nm, box=box, spectral_index=-2.0, rms_disp_cells=0.5,
                                        seed=workloads.SEED_ZELDOVICH, workers=-1, keep_f64=True)
    n = parts.n
    pos = parts.p64
    q = np.stack(np.meshgrid(*(np.arange(n_side) * (box / n_side),) * 3, indexing="ij"), -1)[..., ::-1]
    psi = (pos - q.reshape(-1, 3) + box / 2) % box - box / 2
    vel = a0 * a0 * hubble * psi  # growing mode, f = 1: p = a^2 H psi
    mass = parts.mass64()
    plan = pm.Plan(nm, box, device=0)
    tuning = plan.autotune(reps=3) if not args.no_autotune else plan.tuning()
    lf = gi.Leapfrog(plan, n)
    lf.set_state(pos, vel, mass)
    k = min(args.steps, 32)  # two solves per step; the stage-timing ring holds 64
    a = lf.run(max(args.warmup, 1), dt, 1.0, cosmology=cosmo, a=a0, wrap=True)
    pm.device_synchronize()
    plan.set_timing(True)
    t0 = time.perf_counter()
    a = lf.run(k, dt, 1.0, cosmology=cosmo, a=a, wrap=True)
    pm.device_synchronize()
    wall = time.perf_counter() - t0
    stages = plan.stage_times()
    plan.set_timing(False)
    pm_ms = stages.pop("total")
    ms_step = wall * 1e3 / k
    log(f"[bench] {k} KDK steps: {ms_step:.3f} ms/step, {2 * pm_ms:.3f} ms of it in the two PM solves")
    roofline, _ = roofline_of(stages, n, nm, "zeldovich", SOA64_PARTICLES)

    def host_step(p_, v_, a_, plan_, m32):
        """The reference's GPU path: f64 state on the host, f32 over PCIe per force call."""
        g = io.gravity_coupling_qksl(1.0, a_)
        f = io.drift_kick_factors(c_tuple, a_, dt)

        def force(pp):
            p32 = pp.astype(np.float32)
            ax, ay, az = plan_.accelerations_f32(p32[:, 0].copy(), p32[:, 1].copy(), p32[:, 2].copy(), m32, g)
            return np.stack([ax, ay, az], axis=1).astype(np.float64)
        p_, v_, _ = io.kdk_step(p_, v_, f, force)
        return io.wrap_position(p_, box), v_, io.advance_a(c_tuple, a_, dt)

    p_h, v_h, _ = lf.state()
    m32 = mass.astype(np.float32)
    host_step(p_h, v_h, a, plan, m32)
    t1 = time.perf_counter()
    host_step(p_h, v_h, a, plan, m32)
    host_ms = (time.perf_counter() - t1) * 1e3
    log(f"[bench] host-driven step (PCIe per force call, host kick/drift): {host_ms:.1f} ms")

    small = {}
    for grid in (32, 64):
        qd = workloads.ap08_curve(grid)
        gp = pm.Plan(grid, qd.box, device=0)
        pv = qd.pos64() % qd.box
        vv = np.zeros_like(pv)
        mv = qd.mass64()
        sl = gi.Leapfrog(gp, qd.n)
        sl.set_state(pv, vv, mv)
        sl.run(3, 0.01, 1.0, wrap=True)
        pm.device_synchronize()
        t1 = time.perf_counter()
        sl.run(50, 0.01, 1.0, wrap=True)
        pm.device_synchronize()
        dev_ms = (time.perf_counter() - t1) * 1e3 / 50
        m32s = mv.astype(np.float32)

        def hforce(pp, g=1.0):
            p32 = pp.astype(np.float32)
            ax, ay, az = gp.accelerations_f32(p32[:, 0].copy(), p32[:, 1].copy(), p32[:, 2].copy(), m32s, g)
            return np.stack([ax, ay, az], axis=1).astype(np.float64)
        io.run(pv, vv, 2, 0.01, 1.0, hforce, box=qd.box)
        t1 = time.perf_counter()
        io.run(pv, vv, 20, 0.01, 1.0, hforce, box=qd.box)
        h_ms = (time.perf_counter() - t1) * 1e3 / 20
        small[f"grid{grid}"] = {"n": qd.n, "device_resident_ms_per_step": round(dev_ms, 4),
                                "host_driven_ms_per_step": round(h_ms, 4)}
        sl.close()
        gp.close()
    log(f"[bench] AP-08 sizes: {small}")

    result = {
        "metric": "M particle-steps/s, device-resident KDK leapfrog (2 PM forces per step), 256^3 Zel'dovich "
                  "particles on a 512^3 mesh, LambdaCDM",
        "value": round(n * k / wall / 1e6, 2), "unit": "M particle-steps/s", "n_gpus": 1, "steps": k,
        "warmup": max(args.warmup, 1), "ms_per_step": round(ms_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 state, f32 force", "data": "synthetic",
        "config": {"workload": f"Zel'dovich lattice {n_side}^3 on {nm}^3 mesh, L={box:g}, growing-mode momenta "
                               f"p = a^2 H psi, Omega_m=0.3, Omega_L=0.7, h0={h0:.4f} (G=1, mean density 1), "
                               f"a0={a0}, dt={dt:.4e} (1% growth of a), wrap on",
                   "particles": n, "mesh": nm, "launch_choices": tuning, "final_a": a},
        "roofline": roofline,
        "pm_ms_per_solve": round(pm_ms, 4),
        "integrator_ms_per_step": round(ms_step - 2 * pm_ms, 4),
        # kick+drift and kick+wrap move 120 bytes per particle each
        "integrator_gbps": round(240.0 * n / ((ms_step - 2 * pm_ms) * 1e-3) / 1e9, 1),
        "stage_ms": {kk: round(v, 5) for kk, v in stages.items()},
        "host_driven_ms_per_step": round(host_ms, 2),
        "device_resident_speedup": round(host_ms / ms_step, 2),
        "ap08_sizes": small,
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        import oracle  # CPU baseline only

        oracle.build()
        p_c, v_c, _ = lf.state()
        t1 = time.perf_counter()
        io.run(p_c, v_c, 1, dt, 1.0, lambda pp, g: oracle.pm_accelerations(pp, mass, box, nm, g,
                                                                           mode=oracle.MODE_LEAN),
               cosmology=c_tuple, a=a, box=box)
        tc = time.perf_counter() - t1
        result["cpu_baseline"] = {"value": round(n / tc / 1e6, 4), "unit": "M particle-steps/s",
                                  "cores": oracle.get_threads(), "kind": "port",
                                  "sample": f"one KDK step of the full state (2 FP64 oracle PM solves, lean R2C "
                                            f"mode, plus the f64 kick/drift in numpy): {tc:.2f} s"}
    lf.close()
    plan.close()
    print(json.dumps(result), flush=True)
    return 0


def bench_pk(args) -> int:
    """Power spectrum of the headline particles (256^3 Zel'dovich lattice on a
    512^3 mesh): deposit, R2C x and y passes, z transform with the shell sums,
    on device-resident particles; checked against and timed beside the FP64
    oracle restatement of the reference estimator."""
    import numpy as np

    import gadget_ng_pm_hip as pm
    from gadget_ng_pm_hip import analysis, workloads

    nm, n_side = args.mesh, args.n_side
    box = float(n_side)
    parts = workloads.zeldovich_lattice(n_side=n_side, nm=nm, box=box, spectral_index=-2.0, rms_disp_cells=0.5,
                                        seed=workloads.SEED_ZELDOVICH, workers=-1)
    n = parts.n
    d = [pm.DeviceArray.from_host(a) for a in (parts.x, parts.y, parts.z, parts.m)]
    plan = pm.Plan(nm, box, device=0)
    tuning = plan.autotune(reps=3) if not args.no_autotune else plan.tuning()

    def est():
        return analysis.power_spectrum_device(plan, d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, n)

    for _ in range(max(args.warmup, 1)):
        est()
    k = min(args.steps, 64)
    plan.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(k):
        got = est()
    wall = time.perf_counter() - t0
    stages = plan.stage_times()
    plan.set_timing(False)
    stages.pop("total")
    stages = {("fft_z_shells" if kk == "fft_z_green" else kk): v for kk, v in stages.items()
              if kk not in ("fft_y_inv", "fft_x_inv", "interp")}
    ms = wall * 1e3 / k
    log(f"[bench] P(k): {ms:.3f} ms per estimate (synchronous, shells copied back); stage ms: "
        + ", ".join(f"{kk}={v:.4f}" for kk, v in stages.items()))
    sb = stage_bytes(n, nm)
    sb["fft_z_shells"] = 4 * nm ** 3  # the half spectrum read once
    merge_chunked(stages, sb, {"chunk_fwd": tuning.get("chunk_fwd", 0),
                               "y_fwd_persistent": tuning.get("y_fwd_persistent", False)})
    dom = max(stages, key=stages.get)
    achieved = sb[dom] / (stages[dom] * 1e-3) / 1e9
    result = {
        "metric": "M particles/s through the power-spectrum estimator (CIC deposit, R2C FFT, deconvolved "
                  "shell sums), 256^3 Zel'dovich particles on a 512^3 mesh",
        "value": round(n * k / wall / 1e6, 2), "unit": "M particles/s", "n_gpus": 1, "steps": k,
        "warmup": max(args.warmup, 1), "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 transform, f64 shell sums", "data": "synthetic",
        "config": {"workload": f"Zel'dovich lattice {n_side}^3 particles on {nm}^3 mesh, L={box:g} Mpc/h, "
                               f"P(k)~k^-2, rms displacement 0.5 cell, seed 0x6E670002",
                   "particles": n, "mesh": nm, "launch_choices": tuning},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None,
                     "algorithmic_bytes": sb[dom], "mean_ms": round(stages[dom], 5)},
        "stage_ms": {kk: round(v, 5) for kk, v in stages.items()},
        "shells": int(len(got[0])),
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import analysis_oracle as ao  # checker and CPU baseline only

        t1 = time.perf_counter()
        ref = ao.power_spectrum(parts.pos64(), parts.mass64(), box, nm)
        tc = time.perf_counter() - t1
        result["cpu_baseline"] = {"value": round(n / tc / 1e6, 4), "unit": "M particles/s", "cores": 1,
                                  "kind": "port",
                                  "sample": f"the full workload through the FP64 numpy restatement (CIC by "
                                            f"bincount, complex fftn, shell sums), one run: {tc:.1f} s"}
        err = np.abs(got[1] - ref[1]) / np.maximum(ref[1], 1e-6 * ref[1].max())
        result["modes_identical"] = bool(np.array_equal(got[2], ref[2]))
        result["max_shell_rel_err"] = float(f"{err.max():.3e}")
        log(f"[bench] FP64 oracle {tc:.1f} s; max shell error {err.max():.2e}")
    plan.close()
    print(json.dumps(result), flush=True)
    return 0


def bench_ic(args) -> int:
    """Zel'dovich initial conditions of --mesh^3 particles from a --mesh^3
    Gaussian field (the reference's generator uses one side for both), into
    device memory: per-mode ChaCha12 Gaussians in the z pass, the inverse y and
    x passes of the force pipeline, the lattice kernel.  Timed beside the
    FP64 numpy restatement of the reference generator on a 64^3 sample."""
    import numpy as np

    import gadget_ng_pm_hip as pm
    from gadget_ng_pm_hip import ic
    from gadget_ng_pm_hip import integrate as gi

    n = args.mesh
    box = float(n)
    seed = 0x6E670006
    cosmo = gi.CosmologyParams(0.3, 0.7, 0.1)
    a_init = 0.02
    plan = pm.Plan(n, box, device=0)
    fn = ic.power_law_sigma(n, 1.0e3, -2.0)
    sig = ic.sigma_table(n, fn)
    vf, sc, _ = ic.momentum_factors(cosmo, a_init, 0.1, True)
    npart = n ** 3
    dpos = pm.DeviceArray((npart, 3), np.float64)
    dvel = pm.DeviceArray((npart, 3), np.float64)

    def gen():
        ic.zeldovich_ics_device(plan, seed, sig, vf, sc, dpos.ptr, dvel.ptr)

    for _ in range(max(args.warmup, 1)):
        gen()
    pm.device_synchronize()
    k = args.steps
    t0 = time.perf_counter()
    for _ in range(k):
        gen()
    pm.device_synchronize()
    wall = time.perf_counter() - t0
    ms = wall * 1e3 / k
    # second order (2LPT) into the same buffers, timed the same way
    v1, v2, d2, sc2 = ic.factors_2lpt(cosmo, a_init, 0.1, True)
    ic.zeldovich_2lpt_ics_device(plan, seed, sig, v1, v2, d2, sc2, dpos.ptr, dvel.ptr)
    k2 = max(1, min(k, 5))
    t2 = time.perf_counter()
    for _ in range(k2):
        ic.zeldovich_2lpt_ics_device(plan, seed, sig, v1, v2, d2, sc2, dpos.ptr, dvel.ptr)
    pm.device_synchronize()
    ms2 = (time.perf_counter() - t2) * 1e3 / k2
    log(f"[bench] 2LPT ICs: {ms2:.3f} ms for {npart} particles")
    mh = n * n * (n // 2)
    # compulsory HBM bytes: the z pass writes A and B (8 B each per half-spectrum
    # mode); the y pass reads them and writes three fields; the x pass reads
    # three complex and writes three real fields; the lattice kernel reads the
    # three real fields and writes 48 B per particle
    total_bytes = 16 * mh + 40 * mh + 24 * mh + 12 * n ** 3 + 12 * n ** 3 + 48 * npart
    log(f"[bench] ICs: {ms:.3f} ms for {npart} particles ({npart / (wall / k) / 1e6:.1f} M particles/s)")
    result = {
        "metric": f"M particles/s of Zel'dovich initial conditions ({n}^3 lattice from a {n}^3 Gaussian field: "
                  "per-mode ChaCha12 Gaussians, three inverse transforms, lattice displacement), device-resident",
        "value": round(npart * k / wall / 1e6, 2), "unit": "M particles/s", "n_gpus": 1, "steps": k,
        "warmup": max(args.warmup, 1), "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64 modes and particles, f32 transforms", "data": "synthetic",
        "config": {"workload": f"Zel'dovich ICs {n}^3 particles, L={box:g} Mpc/h, power law A=1e3 n_s=-2, "
                               f"LambdaCDM a_init={a_init}, rescaled to D(1), seed {seed:#x}",
                   "particles": npart, "mesh": n},
        "roofline": {"bound": "hbm", "kernel": "whole generator (the z pass is integer-ALU bound)",
                     "achieved": round(total_bytes / (ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(total_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4), "traffic": None,
                     "algorithmic_bytes": total_bytes},
        "second_order_ms_per_step": round(ms2, 4),
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import ic_oracle as ico  # checker and CPU baseline only

        ns = 64
        sig_s = ic.sigma_table(ns, ic.power_law_sigma(ns, 1.0e3, -2.0))
        t1 = time.perf_counter()
        d = ico.delta_kspace(ns, seed, lambda q: sig_s[np.asarray(q)])
        psi = ico.displacement(d, float(ns))
        rp, rv = ico.zeldovich(ns, float(ns), psi, vf, sc)
        tc = time.perf_counter() - t1
        result["cpu_baseline"] = {"value": round(ns ** 3 / tc / 1e6, 4), "unit": "M particles/s", "cores": 1,
                                  "kind": "port",
                                  "sample": f"64^3 particles through the FP64 numpy restatement of the reference "
                                            f"generator (vectorised ChaCha12, complex ifftn), one run: {tc:.2f} s"}
        sp = pm.Plan(ns, float(ns), device=0)
        gp, gv, _ = ic.zeldovich_ics(sp, seed, ic.power_law_sigma(ns, 1.0e3, -2.0), cosmology=cosmo, a_init=a_init,
                                     rescale_to_a_init=True)
        err = float(np.sqrt(((gv - rv) ** 2).sum() / (rv ** 2).sum()))
        result["sample_rms_rel_err"] = float(f"{err:.3e}")
        log(f"[bench] FP64 oracle sample {tc:.2f} s; momenta rms rel err {err:.2e}")
        sp.close()
    plan.close()
    print(json.dumps(result), flush=True)
    return 0


BF16_DENSE_PEAK_TFLOPS = 2500.0
FP32_PEAK_TFLOPS = 157.3
# FP32 operations of every evaluated short-range pair before the cut: 3
# differences and the squared distance (a multiply and two fused adds, 5
# flops); the periodic offset is applied once per neighbour cell, not per pair
SR_FLOP_PER_PAIR_MIN = 8
NET_FLOP_PER_CELL = 2 * 27 * (8 * 16 + 16 * 16 + 16 * 3)  # 23 328
# compulsory HBM bytes per cell of one inference: x0 in (8 bf16), x1 out and
# in (16 bf16 each way), x2 out and in, y out (4 f32)
NET_BYTES_PER_CELL = 16 + 32 + 32 + 32 + 32 + 16  # 160
# 23 328 / 160 = 146 flop/byte, below the bf16 ridge of 2500 TFLOP/s / 8 TB/s = 312 flop/byte:
# the roofline that bounds the network is HBM bandwidth


# force-error figures of the learned bench: field = particles in cells at most
# (1 + dense_contrast) times the mean mass; field_near = field particles in
# cells next to a denser one, field_far = the rest of the field
ERROR_KEYS = ("field_rms_rel", "field_median_rel", "field_mean_rel", "field_near_rms_rel", "field_far_rms_rel",
              "field_far_median_rel", "rms_rel", "median_rel", "dense_median_rel")


def bench_learned(args) -> int:
    """Learned sub-cell correction (clustered boxes at one particle per cell): train on seeded 64^3
    boxes made by the method itself, evaluate on held-out 256^3 boxes against
    TreePM (long range + exact short range), and time the network against the
    short-range sum it stands in for."""
    import numpy as np

    import gadget_ng_pm_hip as pm
    from gadget_ng_pm_hip import learned

    if args.copy_calibration:
        src = pm.DeviceArray((1 << 28,), np.float32)
        dst = pm.DeviceArray((1 << 28,), np.float32)
        pm.device_copy(dst, src)
        pm.device_synchronize()
        del src, dst
    lc = learned.LearnedCorrection(cells=64, batch=8)
    t0 = time.perf_counter()
    info = lc.build_training_set(range(1, args.train_boxes + 1), workers=-1)
    t_data = time.perf_counter() - t0
    log(f"[learned] {args.train_boxes} training boxes in {t_data:.1f} s: {info}")
    t0 = time.perf_counter()
    hist = lc.train(steps=args.train_steps, lr=2e-3, log_every=max(1, args.train_steps // 10),
                    log=lambda m: log("[learned] " + m))
    t_train = time.perf_counter() - t0
    evals = [lc.evaluate(1001 + i, cells=256, workers=-1) for i in range(args.eval_boxes)]
    for e in evals:
        log(f"[learned] seed {e['seed']}: PM field rms {e['pm']['field_rms_rel']:.4f} -> "
            f"{e['pm_net']['field_rms_rel']:.4f}, field median {e['pm']['field_median_rel']:.5f} -> "
            f"{e['pm_net']['field_median_rel']:.5f}")
    # network inference roofline at 256^3: repeated forwards on the last evaluation's input
    big = learned.LearnedCorrection(cells=256, batch=1)
    big.net.set_params(lc.net.params())
    pm.device_synchronize()
    big.net.infer(1)
    pm.device_synchronize()
    reps = 10
    t0 = time.perf_counter()
    for _ in range(reps):
        big.net.infer(1)
    pm.device_synchronize()
    t_fwd = (time.perf_counter() - t0) / reps
    cells = 256 ** 3
    achieved = NET_FLOP_PER_CELL * cells / t_fwd / 1e12

    def mean(key, which):
        return float(np.mean([e[which][key] for e in evals]))

    t_sr = float(np.mean([e["time_short_range_s"] for e in evals]))
    sr_pairs = float(np.mean([e["sr_pairs"] for e in evals]))
    t_net = float(np.mean([e["time_net_s"] for e in evals]))
    result = {
        "metric": "learned sub-cell correction, 256^3 clustered particles on a 256^3 mesh: held-out field-particle "
                  "rms force error vs TreePM",
        "value": round(mean("field_rms_rel", "pm_net"), 5),
        "unit": "relative rms force error",
        "n_gpus": 1, "steps": args.train_steps, "warmup": 0, "ms_per_step": round(t_train * 1e3 / args.train_steps, 4),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "bf16", "data": "synthetic",
        "config": {"workload": "clustered boxes: 70% Zel'dovich lattice + 30% in Plummer spheres (a = 0.2 Mpc/h), one "
                               "particle per (Mpc/h)^3; train on 64^3 boxes (seeds 1..N), evaluate on 256^3 boxes "
                               "(seeds 1001..)",
                   "train_boxes": args.train_boxes, "eval_boxes": args.eval_boxes,
                   "treepm": "r_s = 1.25 cells, CIC-deconvolved long range + exact short range within 5 r_s, "
                             "eps = 0.05"},
        "target": {"rms_rel": 0.01, "met_field": bool(mean("field_rms_rel", "pm_net") <= 0.01),
                   "met_all": bool(mean("rms_rel", "pm_net") <= 0.01)},
        "without_net": {k: round(mean(k, "pm"), 5) for k in ERROR_KEYS},
        "with_net": {k: round(mean(k, "pm_net"), 5) for k in ERROR_KEYS},
        "populations": {k: int(round(mean(k, "pm"))) for k in ("field_count", "field_near_count", "field_far_count",
                                                                "dense_count")},
        "cost_s": {"network_forward_and_apply": round(t_net, 5), "short_range_sum": round(t_sr, 5),
                   "network_input": round(float(np.mean([e["time_input_s"] for e in evals])), 5),
                   "pm": round(float(np.mean([e["time_pm_s"] for e in evals])), 5),
                   "treepm_long_range": round(float(np.mean([e["time_treepm_long_s"] for e in evals])), 5)},
        "roofline": {"bound": "hbm", "kernel": "network forward (3 convolutions)",
                     "achieved": round(NET_BYTES_PER_CELL * cells / t_fwd / 1e9, 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(NET_BYTES_PER_CELL * cells / t_fwd / 1e9 / HBM_PEAK_GBPS, 4),
                     "traffic": None, "bytes_per_cell": NET_BYTES_PER_CELL, "mean_ms": round(t_fwd * 1e3, 4),
                     "mfma_tflops": round(achieved, 2), "mfma_frac": round(achieved / BF16_DENSE_PEAK_TFLOPS, 5),
                     "flop_per_cell": NET_FLOP_PER_CELL},
        # the short-range direct sum (k_force): every particle against every
        # particle of its 27 chaining-mesh cells; each evaluated pair costs at
        # least SR_FLOP_PER_PAIR_MIN FP32 operations (differences, squared
        # distance) before the cut, so this rate bounds the kernel's
        # FP32 throughput from below
        "short_range": {"ms": round(t_sr * 1e3, 2), "pairs_evaluated": int(sr_pairs),
                        "gpairs_per_s": round(sr_pairs / t_sr / 1e9, 1),
                        "flop_per_pair_min": SR_FLOP_PER_PAIR_MIN,
                        "fp32_tflops_min": round(sr_pairs * SR_FLOP_PER_PAIR_MIN / t_sr / 1e12, 2),
                        "fp32_peak_tflops": FP32_PEAK_TFLOPS,
                        "frac_min": round(sr_pairs * SR_FLOP_PER_PAIR_MIN / t_sr / 1e12 / FP32_PEAK_TFLOPS, 4)},
        "training": {"data_s": round(t_data, 2), "train_s": round(t_train, 2), "loss_first": hist[0],
                     "loss_last": hist[-1], "f0": lc.f0},
    }
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
