#!/usr/bin/env python3
"""Benchmark of the HOS time-step path on one MI355X.

Workload (the single-instance 1024^2 configuration): 1024x1024 grid over a
32x32 peak-wavelength square domain, kp = 1 m^-1, finite depth kp*h = 1.0,
HOS order M = 5 (nlevs = 4), JONSWAP gamma = 5, kp*Hs/2 = 0.12, cos^2
spreading, |k| <= 4 kp, seed 1, classical RK4 with dt = Tp/100.

A step is one RK4 step (4 RHS evaluations) of every member.  Timed in order:
  1. the FP32 batched form: --members (default 8) realisations of the
     workload advanced together (FFTs, spectral multiplies and products in
     FP32 with Kahan-compensated order sums; FP64 state, stage combination,
     linear terms and diagnostics), launch shapes chosen by the autotuner
     (y tiles, x-pass mode, one chain or two member groups on their own
     streams), captured graphs;
  2. the same FP32 form on one instance;
  3. the plain FP64 form on one instance, same input.
value = grid-point RK4 steps/s of (1): members * Nx * Ny * steps / wall time
of the timed steps, state resident in HBM, between two device
synchronisations.  p50/p99 are per-step times from HIP events.
roofline: a separate stream-ordered pass of the same batch (one chain, so
kernels do not overlap) with a HIP event pair around every kernel on the
solver's stream; the dominant kernel's algorithmic bytes per launch over its
mean duration.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth


def workload():
    kp = 1.0
    L = 32 * 2 * math.pi / kp
    return dict(Nx=1024, Ny=1024, Lx=L, Ly=L, kp=kp, depth=1.0 / kp, nlevs=4, gamma=5.0, eps=0.12, seed=1,
                g=9.81)


def make_state(hos, w):
    return hos.jonswap(w["Nx"], w["Ny"], w["Lx"], w["Ly"], kp=w["kp"], gamma=w["gamma"], eps=w["eps"],
                       depth=w["depth"], g=w["g"], seed=w["seed"], nlevs=w["nlevs"])


def run_form(hos, w, prec, members, steps, warmup, dt, u0, graph=True, profile_steps=0, autotune=True, tiles=None):
    p = hos.params(Nx=w["Nx"], Ny=w["Ny"], nlevs=w["nlevs"], Lx=w["Lx"], Ly=w["Ly"], g=w["g"],
                   depth=w["depth"], precision=prec, compensated=1)
    s = hos.Solver(p, members)
    # u0: one state for every member, or a list of one state per member
    for m in range(members):
        s.set_state(u0[m] if isinstance(u0, list) else u0, m)
        s.set_member(m, w["depth"], dt)
    if tiles:
        # fixed launch shapes (c2r, r2c, pair y tiles; x-pass mode), no tuning trials
        for kind, v in enumerate(tiles[:3]):
            s.set_tile(v, kind)
        if len(tiles) > 3:
            s.set_tile(tiles[3], hos.HOS_TILE_XPASS)
        if len(tiles) > 4:
            s.set_tile(tiles[4], hos.HOS_TILE_GROUPS)
        if len(tiles) > 5:
            s.set_tile(tiles[5], hos.HOS_TILE_PLAN)
        if len(tiles) > 6:
            s.set_tile(tiles[6], hos.HOS_TILE_FUSE)
        if len(tiles) > 7:
            s.set_tile(tiles[7], hos.HOS_TILE_SKEW)
        if len(tiles) > 8:
            s.set_tile(tiles[8], hos.HOS_TILE_DIAG)
        if len(tiles) > 9:
            s.set_tile(tiles[9], hos.HOS_TILE_XBLOCK)
        tiles = tuple(tiles[:3])
    else:
        tiles = s.autotune(3) if autotune else None
    xpass = s.get_tile(hos.HOS_TILE_XPASS)
    groups = s.get_tile(hos.HOS_TILE_GROUPS)
    plan = s.get_tile(hos.HOS_TILE_PLAN)
    fuse = s.get_tile(hos.HOS_TILE_FUSE)
    skew = s.get_tile(hos.HOS_TILE_SKEW)
    diag = s.get_tile(hos.HOS_TILE_DIAG)
    xblock = s.get_tile(hos.HOS_TILE_XBLOCK)
    s.use_graph(graph)
    s.step(warmup, dt)
    s.sync()
    t0 = time.perf_counter()
    s.step(steps, dt)  # returns after the device has finished the K steps
    s.sync()
    wall = time.perf_counter() - t0
    st = s.step_times()
    out = {
        "wall_s": wall,
        "ms_per_step": 1e3 * wall / steps,
        "value": members * w["Nx"] * w["Ny"] * steps / wall,
        "p50_ms": float(np.percentile(st, 50)),
        "p99_ms": float(np.percentile(st, 99)),
        "H": s.energy(0)[0],
        "vol": s.energy(0)[1],
        "state": s.get_state(0),
        "bytes_per_pt_step": hos.bytes_per_point_step(p),
        "tiles": tiles,
        "xpass": "row per wave" if xpass == 2 else "rows per workgroup",
        "groups": groups,
        "plan": plan,
        "fuse": fuse == 2 and diag != 2,
        "skew": skew == 2 and groups > 1,
        "diag": diag == 2,
        "xblock": xblock == 2 and diag == 2 and xpass == 2 and p.nlevs >= 1,
    }
    if profile_steps:
        # same workload, stream-ordered launches with a HIP event pair around
        # every kernel on the solver's stream
        s.use_graph(False)
        s.profile(True)
        s.step(profile_steps, dt)
        out["kstats"] = s.kernel_stats()
        out["profile_ms_per_step"] = s.last_step_ms()
        s.profile(False)
    s.close()
    return out


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """Host threads this process may use: its CPU affinity, capped by
    OMP_NUM_THREADS when set (the share of the host that goes with one GPU)."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(n, 1)


def cpu_baseline(w, dt, u0, steps=4):
    """The FP64 CPU restatement of the reference path (the oracle: the
    reference's Mult structure, 3 FFTs per product) on the same workload: one
    RK4 step on one thread, then `steps` RK4 steps with the transforms on all
    the host threads this process may use (OpenMP), each step timed
    (p50/p99).  The reference FFTW-MPI build needs FFTW and MPI, which this
    image does not have."""
    import oracle

    o = oracle.Oracle(w["Nx"], w["Ny"], w["nlevs"], w["Lx"], w["Ly"], g=w["g"], depth=w["depth"], scheme=2)
    u = np.ascontiguousarray(u0.copy())
    oracle.set_threads(1)
    t0 = time.perf_counter()
    o.step(u, 0.0, dt)
    sec1 = time.perf_counter() - t0
    P = oracle.set_threads(cpu_threads())
    t = dt
    per = []
    for _ in range(steps):
        t0 = time.perf_counter()
        t, _ = o.step(u, t, dt)
        per.append(time.perf_counter() - t0)
    oracle.set_threads(1)
    pts = w["Nx"] * w["Ny"]
    return {
        "value": pts * steps / sum(per),
        "unit": "grid-pt RK4 steps/s",
        "cores": P,
        "kind": "port",
        "sample": f"{steps} RK4 steps of the same 1024^2 M=5 workload on {P} threads ({sum(per):.1f} s) after "
                  f"1 step on 1 thread ({sec1:.1f} s); FP64 CPU restatement with the reference's Mult structure "
                  "(3 FFTs per product), own radix-2 FFT, OpenMP over the row transforms; the reference "
                  "FFTW-MPI build is not buildable in this image (no FFTW, no MPI)",
        "p50_ms": round(1e3 * float(np.percentile(per, 50)), 1),
        "p99_ms": round(1e3 * float(np.percentile(per, 99)), 1),
        "single_thread_value": pts / sec1,
        "cpu_model": cpu_model(),
    }


def other_workloads(serve_seconds):
    """The ensemble, uneven-batch and serving workloads (hos_amd.serve_loop),
    each timed on its own: 64 realisations of 512^2 at M=4 for 500 RK4 steps;
    96 uneven requests for 300 steps; Poisson arrivals at 20 requests/s for
    serve_seconds (the 120 s of the full serving workload is the same load
    for longer)."""
    from hos_amd import serve_loop as sl

    out = {}
    b = sl.batch(members=64, N=512, nlevs=3, steps=500, check=2)
    out["ensemble_64x512_m4"] = b
    print(f"[bench] ensemble 64 x 512^2: {b['value']:.3e} grid-pt RK4 steps/s", file=sys.stderr, flush=True)
    e = sl.ensemble(n=96, steps=300)
    out["uneven_96x300"] = e
    print(f"[bench] uneven 96 x 300: {e['value']:.3e} grid-pt RK4 steps/s", file=sys.stderr, flush=True)
    if serve_seconds > 0:
        v = sl.serve(rate=20.0, duration=serve_seconds)
        out["serve_poisson_20rps"] = v
        print(f"[bench] serve {serve_seconds:.0f} s: {v['value']:.3e} grid-pt RK4 steps/s", file=sys.stderr,
              flush=True)
    return out


# kernel name -> profiled kind (hos_kernel_stats), first match: the RHS pass
# k_yr2c_rhs is the "yr2c_pair" kind (it replaced the plain pair pass in the
# step); k_yc2r_phi before its prefix k_yc2r
KIND_OF = [("k_yr2c_pair", "yr2c_pair"), ("k_yr2c_rhs", "yr2c_pair"), ("k_yr2c", "yr2c"), ("k_yc2r_phi", "yc2r_phi"),
           ("k_ydiag", "ydiag"),
           ("k_yc2r", "yc2r"), ("k_xfwdinv", "xfwdinv"), ("k_xfwd_rhs", "xfwd_rhs"), ("k_xinv", "xinv"),
           ("k_copy_c64", "copy")]
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_1024_m5_x8.json")


def traffic_summary(fetch_dir, write_dir, cal_fetch_dir, cal_write_dir):
    """HBM bytes per launch per kernel kind from rocprofv3 --pmc FETCH_SIZE and
    WRITE_SIZE runs (KiB per dispatch), each scaled by what the counter reads
    for the plain copy kernel of known size (bench.py --calib)."""
    import csv
    from collections import defaultdict

    def per_kind(d):
        acc = defaultdict(list)
        for r in csv.DictReader(open(os.path.join(d, "run_counter_collection.csv"))):
            name = r["Kernel_Name"]
            kind = next((k for pat, k in KIND_OF if pat in name), None)
            if kind:
                acc[kind].append(float(r["Counter_Value"]))
        return {k: sum(v) / len(v) for k, v in acc.items()}, {k: len(v) for k, v in acc.items()}

    f, nf = per_kind(fetch_dir)
    w, _ = per_kind(write_dir)
    cf, _ = per_kind(cal_fetch_dir)
    cw, _ = per_kind(cal_write_dir)
    copy_bytes = 256 << 20  # read and written once per copy dispatch
    sf = copy_bytes / (cf["copy"] * 1024.0)
    sw = copy_bytes / (cw["copy"] * 1024.0)
    out = {"fetch_scale": sf, "write_scale": sw, "kernels": {}}
    for k in f:
        if k == "copy" or k not in w:
            continue
        out["kernels"][k] = {"dispatches": nf[k], "fetch_KiB": f[k], "write_KiB": w[k],
                             "bytes_per_launch": (f[k] * sf + w[k] * sw) * 1024.0}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--members", type=int, default=8, help="1024^2 realisations in the FP32 batched form")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-single", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=5)
    ap.add_argument("--no-others", action="store_true",
                    help="skip the ensemble, uneven-batch and serving workloads")
    ap.add_argument("--serve-seconds", type=float, default=10.0,
                    help="arrival window of the Poisson serving workload (0: skip it)")
    ap.add_argument("--tiles", type=str, default=None,
                    help="fixed launch shapes c2r,r2c,pair[,xpass[,groups[,plan[,fuse[,skew[,diag[,xblock]]]]]]] for the batched form instead of the autotuner "
                         "(profiler runs: no tuning trials in the kernel statistics)")
    ap.add_argument("--calib", action="store_true", help="only the plain copy kernel (counter calibration)")
    ap.add_argument("--traffic", nargs=4, metavar=("FETCH", "WRITE", "CAL_FETCH", "CAL_WRITE"),
                    help="summarise four rocprofv3 --pmc output folders into --traffic-out")
    ap.add_argument("--traffic-out", default=TRAFFIC_FILE, help="file the --traffic summary goes to")
    ap.add_argument("--precision", choices=("fp32", "fp64"), default="fp32",
                    help="arithmetic of the batched form (fp64: the FP64 form in its place, for profiling runs "
                         "of the FP64 kernels with --members 1 --no-fp64)")
    args = ap.parse_args()
    if args.traffic:
        t = traffic_summary(*args.traffic)
        with open(args.traffic_out, "w") as fh:
            json.dump(t, fh, indent=1)
        print(json.dumps(t))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); the members of an "
              "ensemble would shard across GPUs with no communication between them", file=sys.stderr)
        sys.exit(2)

    import hos_amd as hos

    if args.calib:
        # 256 MiB copied 10 times: the per-dispatch counter reading of a known
        # 512 MiB (read + write) of traffic
        gbps = hos.copy_bandwidth(256 << 20, 10)
        print(json.dumps({"copy_bytes_per_dispatch": 2 * (256 << 20), "copy_GBps": round(gbps, 1)}))
        return

    w = workload()
    Tp = hos.peak_period(w["g"], w["kp"], w["depth"])
    dt = Tp / 100.0
    u0 = make_state(hos, w)
    # the batch: one sea state per member, seeds 1 .. members (member 0 is the
    # seed-1 state the single-instance forms step)
    batch = [u0] + [make_state(hos, dict(w, seed=w["seed"] + m)) for m in range(1, args.members)]
    print(f"[bench] device: {hos.device_name()}", file=sys.stderr, flush=True)
    graph = not args.no_graph

    fixed = [int(v) for v in args.tiles.split(",")] if args.tiles else None
    prec = hos.HOS_FP32 if args.precision == "fp32" else hos.HOS_FP64
    f32 = run_form(hos, w, prec, args.members, args.steps, args.warmup, dt, batch, graph=graph,
                   profile_steps=args.profile_steps, tiles=fixed)
    print(f"[bench] {args.precision} x{args.members}: {f32['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
    f1 = None
    if not args.no_single and args.members != 1:
        f1 = run_form(hos, w, hos.HOS_FP32, 1, args.steps, args.warmup, dt, u0, graph=graph)
        print(f"[bench] fp32 x1: {f1['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
    f64 = None
    if not args.no_fp64:
        f64 = run_form(hos, w, hos.HOS_FP64, 1, args.steps, args.warmup, dt, u0, graph=graph)
        print(f"[bench] fp64 x1: {f64['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
    copy_gbps = hos.copy_bandwidth(256 << 20, 10)

    # dominant kernel: the kernel (one template instance) with the largest
    # summed duration in the profiled pass; the fused level passes are one
    # instance per level 2 .. M-1, the diagonal passes one per diagonal 1 .. M-1
    ks = f32.get("kstats", {})
    roof = None
    if ks:
        inst = {n: (w["nlevs"] - 1 if n == "yc2r_phi" else (w["nlevs"] if n == "ydiag" else 1)) for n in ks}
        name = max(ks, key=lambda k: ks[k]["ms"] / inst[k])
        k = ks[name]
        avg_ms = k["ms"] / max(k["launches"], 1)
        avg_bytes = k["bytes"] / max(k["launches"], 1)
        achieved = avg_bytes / (avg_ms * 1e-3) / 1e9
        traffic = None
        if args.members == 8 and args.precision == "fp32" and os.path.exists(TRAFFIC_FILE):
            tk = json.load(open(TRAFFIC_FILE)).get("kernels", {}).get(name)
            traffic = round(tk["bytes_per_launch"]) if tk else None
        roof = {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": name,
            "avg_launch_us": round(avg_ms * 1e3, 2), "bytes_per_launch": avg_bytes,
            "kernel_share": round(k["ms"] / sum(v["ms"] for v in ks.values()), 3),
            "practical_copy_GBps": round(copy_gbps, 1),
            "per_kernel": {n: {"launches": v["launches"],
                               "avg_us": round(1e3 * v["ms"] / max(v["launches"], 1), 2),
                               "GBps": round(v["bytes"] / max(v["ms"], 1e-12) / 1e6, 1)}
                           for n, v in ks.items() if v["launches"]},
        }
    cpu = None if args.no_cpu else cpu_baseline(w, dt, u0)
    if cpu:
        print(f"[bench] cpu: {cpu['value']:.3e} grid-pt RK4 steps/s on {cpu['cores']} threads", file=sys.stderr,
              flush=True)
    others = {} if args.no_others else other_workloads(args.serve_seconds)

    def line(f):
        return {"value": round(f["value"], 1), "ms_per_step": round(f["ms_per_step"], 4),
                "p50_ms": round(f["p50_ms"], 4), "p99_ms": round(f["p99_ms"], 4)}

    res = {
        "metric": "grid-pt RK4 steps/s, M=5 1024^2, FP32 batched vs FP64 plain; p50/p99 ms",
        "value": round(f32["value"], 1),
        "unit": "grid-pt RK4 steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(f32["ms_per_step"], 4),
        "p50_ms": round(f32["p50_ms"], 4),
        "p99_ms": round(f32["p99_ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic seeded JONSWAP sea states (hos_ic_jonswap), resident in HBM",
        "config": {
            "workload": f"FP32 batched form: {args.members} realisations of the 1024x1024 grid over 32x32 peak "
                        "wavelengths, kp*h=1, M=5, JONSWAP gamma=5 kp*Hs/2=0.12, "
                        + (f"seeds 1..{args.members} (one sea state per member)" if args.members > 1 else "seed 1")
                        + ", RK4 dt=Tp/100, all members in the same launches; compared with the plain FP64 single "
                        "instance (seed 1)",
            "members": args.members, "graph": graph, "y_tiles_c2r_r2c_pair": f32["tiles"],
            "x_inverse_passes": f32["xpass"], "member_groups": f32["groups"],
            # the plan family of the RHS pass's row FFTs (fft_stage_bits); the built-in
            # rule gives 16 8 8 at 512 threads, 16 16 4 at 256 (the FP32 RHS pass over W fields)
            "rhs_fft_plan": {1: "built-in rule", 2: "radix <= 8", 3: "radix 4", 4: "largest radix"}.get(f32["plan"],
                                                                                                    f32["plan"]),
            "w_expansion": "diagonal passes" if f32["diag"] else "level passes",
            "x_intermediates": "blocked: 4-row column-pair chunks" if f32["xblock"] else "row by row",
            "fused_level_passes": f32["fuse"],
            "offset_member_groups": f32["skew"],
            "form": "FP32 FFTs/products with Kahan order sums; FP64 state, RK combination, linear terms",
            "parallelism": "single GPU",
        },
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if f1:
        res["fp32_single"] = line(f1)
    if f64:
        res["fp64_plain"] = line(f64)
        res["speedup_fp32_batched_vs_fp64_plain"] = round(f32["value"] / f64["value"], 3)
        if f1:
            res["speedup_fp32_single_vs_fp64_plain"] = round(f1["value"] / f64["value"], 3)
        d = f32["state"][0] - f64["state"][0]
        res["eta_rel_l2_fp32_vs_fp64"] = float(np.linalg.norm(d) / np.linalg.norm(f64["state"][0]))
        res["steps_compared"] = args.warmup + args.steps
    res["alg_bytes_per_pt_step"] = f32["bytes_per_pt_step"]
    res["alg_GBps"] = round(f32["bytes_per_pt_step"] * f32["value"] / 1e9, 1)
    res.update(others)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
