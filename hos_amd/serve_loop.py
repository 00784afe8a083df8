"""Ensemble and serving runs over hos_server (include/hos_serve.h).

Request mix (the uneven-batch distribution): grid from {128^2, 256^2, 512^2,
1024^2, 2048x512} over (Nx/16) x (Ny/16) peak wavelengths (kp = 1, Nyquist
8 kp), HOS order M from {2..6}, kp*h from {0.8, 1.5, deep}, JONSWAP gamma in
[1, 7] and kp*Hs/2 in [0.04, 0.13], cos^2 spreading, |k| <= 4 kp, phases
seeded by the request index, RK4 with dt = Tp/64.

  python -m hos_amd.serve_loop ensemble [--n 96 --steps 300]
  python -m hos_amd.serve_loop serve [--rate 20 --duration 120 --hmin 200 --hmax 2000]

Each prints one JSON line: throughput in grid-point RK4 steps/s, p50/p99
step and request latency, and the sampled error of the FP32 results against
the plain FP64 path.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import threading
import time

import numpy as np

import hos_amd as hos

GRIDS = [(128, 128), (256, 256), (512, 512), (1024, 1024), (2048, 512)]
DEPTHS = [0.8, 1.5, 0.0]  # kp*h; 0 = deep


def mix_request(i: int, grids=GRIDS, orders=(2, 3, 4, 5, 6), hmin=300, hmax=300):
    """Parameters of request i, drawn from a generator seeded by i."""
    rng = np.random.default_rng(1_000_003 * (i + 1))
    Nx, Ny = grids[rng.integers(len(grids))]
    M = int(orders[rng.integers(len(orders))])
    kph = DEPTHS[rng.integers(len(DEPTHS))]
    gamma = float(rng.uniform(1.0, 7.0))
    eps = float(rng.uniform(0.04, 0.13))
    steps = int(rng.integers(hmin, hmax + 1))
    kp, g = 1.0, 9.81
    lam = 2 * math.pi / kp
    Lx, Ly = Nx / 16 * lam, Ny / 16 * lam
    depth = kph / kp
    Tp = hos.peak_period(g, kp, depth)
    return dict(Nx=Nx, Ny=Ny, nlevs=M - 1, Lx=Lx, Ly=Ly, g=g, depth=depth, gamma=gamma, eps=eps, seed=i,
                dt=Tp / 64, nsteps=steps, kp=kp)


def initial_state(r):
    return hos.jonswap(r["Nx"], r["Ny"], r["Lx"], r["Ly"], kp=r["kp"], gamma=r["gamma"], eps=r["eps"],
                       depth=r["depth"], g=r["g"], seed=r["seed"], nlevs=r["nlevs"])


def submit(server, r, u0):
    return server.submit(u0, r["Nx"], r["Ny"], r["nlevs"], r["Lx"], r["Ly"], r["dt"], r["nsteps"],
                         depth=r["depth"], g=r["g"])


def run_alone(r, u0, precision, budget=250e9):
    """The same request alone in a server of its own (the reference for 'same
    result alone and in the mix': its pool has the same slots, hence the same
    launch shapes and FFT forms, as in the mix; and the FP64 path for the
    sampled error)."""
    srv = hos.Server(hbm_budget=budget, precision=precision)
    rid = submit(srv, r, u0)
    srv.drain(600.0)
    u, _ = srv.result(rid)
    srv.close()
    return u


def drain_reporting(srv, total, label):
    """Drain the server, printing the completed-request count (stderr) after
    every 10 s slice until nothing is queued or running."""
    while not srv.drain_for(10.0):
        print(f"[{label}] {int(srv.stats()['completed'])} / {total} requests done", file=sys.stderr, flush=True)


def eta_rel_err(state, ref):
    """Relative L2 error of eta-hat (component 0 of a [2][Ny/2+1][Nx] state)."""
    return float(np.linalg.norm(state[0] - ref[0]) / np.linalg.norm(ref[0]))


def sampled_error(reqs, results, frac, seed=0):
    rng = np.random.default_rng(seed)
    ids = [i for i in sorted(results) if rng.random() < frac] or sorted(results)[:1]
    errs = []
    for i in ids:
        r, u0 = reqs[i]
        ref = run_alone(r, u0, hos.HOS_FP64)
        errs.append(eta_rel_err(results[i][0], ref))  # results[i] = (state, info)
    return ids, errs


def shape_key(r):
    return (r["Nx"], r["Ny"], r["nlevs"])


def run_batch(reqs, budget, label):
    """Submit every request of reqs {i: (r, u0)} at once and drain; returns
    (wall seconds, results, stats)."""
    srv = hos.Server(hbm_budget=budget, precision=hos.HOS_FP32)
    ids = {}
    t0 = time.perf_counter()
    for i, (r, u0) in reqs.items():
        ids[i] = submit(srv, r, u0)
    drain_reporting(srv, len(reqs), label)
    wall = time.perf_counter() - t0
    results = {i: srv.result(ids[i]) for i in reqs}
    st = srv.stats()
    srv.close()
    return wall, results, st


def ensemble(n=96, steps=300, budget=250e9, check_frac=0.05, by_shape=True):
    """The uneven batch: n requests of the mix, all submitted at once.  With
    by_shape, the same requests also run one shape at a time (each shape's
    requests alone in the server, shapes one after another): the sum of those
    walls over the mixed wall is how much the mix keeps of single-shape
    batches (1.0: nothing lost to mixing)."""
    reqs = {}
    for i in range(n):
        r = mix_request(i, hmin=steps, hmax=steps)
        reqs[i] = (r, initial_state(r))
    wall, results, st = run_batch(reqs, budget, "ensemble")
    pts = sum(r["Nx"] * r["Ny"] * r["nsteps"] for r, _ in reqs.values())
    out = dict(mode="ensemble", requests=n, steps=steps, wall_s=wall, value=pts / wall,
               unit="grid-pt RK4 steps/s", step_p50_ms=st["step_p50_ms"], step_p99_ms=st["step_p99_ms"],
               pools=st["pools"], peak_bytes=st["peak_bytes"])
    if by_shape:
        shapes = sorted({shape_key(r) for r, _ in reqs.values()})
        walls = []
        for k in shapes:
            sub = {i: v for i, v in reqs.items() if shape_key(v[0]) == k}
            w, res, _ = run_batch(sub, budget, f"shape {k}")
            walls.append(w)
            for i in sub:  # the same bits alone in its shape's batch and in the mix
                assert np.array_equal(res[i][0], results[i][0]), i
        out.update(by_shape_wall_s=sum(walls), shapes=len(shapes),
                   mix_vs_by_shape=sum(walls) / wall)
    sid, errs = sampled_error(reqs, results, check_frac)
    out.update(sampled=len(sid), eta_rel_l2_vs_fp64_max=max(errs))
    return out


def serve(rate=20.0, duration=120.0, hmin=200, hmax=2000, budget=250e9, check_frac=0.05, seed=0):
    rng = np.random.default_rng(seed)
    srv = hos.Server(hbm_budget=budget, precision=hos.HOS_FP32)
    srv.start()
    reqs, ids = {}, {}
    t0 = time.perf_counter()
    i = 0
    next_t = rng.exponential(1.0 / rate)
    while True:
        now = time.perf_counter() - t0
        if next_t > duration:
            break
        if now < next_t:
            time.sleep(min(next_t - now, 0.05))
            continue
        r = mix_request(i, hmin=hmin, hmax=hmax)
        u0 = initial_state(r)
        reqs[i] = (r, u0)
        ids[i] = submit(srv, r, u0)
        i += 1
        next_t += rng.exponential(1.0 / rate)
    t_arrivals = time.perf_counter() - t0
    print(f"[serve] {i} requests arrived in {t_arrivals:.1f} s", file=sys.stderr, flush=True)
    drain_reporting(srv, i, "serve")
    wall = time.perf_counter() - t0
    srv.stop()
    results = {k: srv.result(ids[k]) for k in reqs}
    st = srv.stats()
    srv.close()
    pts = sum(r["Nx"] * r["Ny"] * r["nsteps"] for r, _ in reqs.values())
    sid, errs = sampled_error(reqs, results, check_frac)
    return dict(mode="serve", rate=rate, arrivals_s=t_arrivals, requests=len(reqs), wall_s=wall,
                value=pts / wall, unit="grid-pt RK4 steps/s", step_p50_ms=st["step_p50_ms"],
                step_p99_ms=st["step_p99_ms"], request_p50_s=st["req_p50_s"], request_p99_s=st["req_p99_s"],
                queue_wait_p50_s=st["wait_p50_s"], queue_wait_p99_s=st["wait_p99_s"],
                busy_s=st["busy_s"], pools=st["pools"], peak_running=st["peak_running"],
                peak_bytes=st["peak_bytes"], sampled=len(sid), eta_rel_l2_vs_fp64_max=max(errs))


def batch(members=64, N=512, nlevs=3, steps=500, check=2, precision=None):
    """Uniform batch: members realisations on NxN over 16x16 peak wavelengths,
    deep water, M = nlevs+1, JONSWAP gamma = 3.3, kp*Hs/2 uniform in
    [0.06, 0.12], seeds 0..members-1, dt = Tp/64, all members in the same
    launches (one captured graph per step)."""
    kp, g = 1.0, 9.81
    L = 16 * 2 * math.pi / kp
    Tp = hos.peak_period(g, kp, 0.0)
    dt = Tp / 64
    prec = hos.HOS_FP32 if precision is None else precision
    p = hos.params(Nx=N, Ny=N, nlevs=nlevs, Lx=L, Ly=L, g=g, precision=prec, compensated=1)
    s = hos.Solver(p, members)
    u0 = []
    for m in range(members):
        eps = 0.06 + 0.06 * np.random.default_rng(m).random()
        u = hos.jonswap(N, N, L, L, kp=kp, gamma=3.3, eps=eps, seed=m, nlevs=nlevs)
        u0.append(u)
        s.set_state(u, m)
        s.set_member(m, 0.0, dt)
    tiles = s.autotune(2)
    s.use_graph(True)
    s.step(2, dt)  # warm
    for m in range(members):
        s.set_state(u0[m], m)
        s.set_member(m, 0.0, dt)
    t0 = time.perf_counter()
    s.step(steps, dt)
    wall = time.perf_counter() - t0
    st = s.step_times()
    errs = []
    for m in range(min(check, members)):
        ref = run_alone(dict(Nx=N, Ny=N, nlevs=nlevs, Lx=L, Ly=L, g=g, depth=0.0, nsteps=steps, dt=dt), u0[m],
                        hos.HOS_FP64)
        errs.append(float(np.linalg.norm(s.get_state(m)[0] - ref[0]) / np.linalg.norm(ref[0])))
    s.close()
    return dict(mode="batch", members=members, N=N, M=nlevs + 1, steps=steps, wall_s=wall,
                value=members * N * N * steps / wall, unit="grid-pt RK4 steps/s", ms_per_step=1e3 * wall / steps,
                step_p50_ms=float(np.percentile(st, 50)), step_p99_ms=float(np.percentile(st, 99)),
                tiles=list(tiles), eta_rel_l2_vs_fp64=errs)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["ensemble", "serve", "batch"])
    ap.add_argument("--members", type=int, default=64)
    ap.add_argument("--N", type=int, default=512)
    ap.add_argument("--nlevs", type=int, default=3)
    ap.add_argument("--n", type=int, default=96)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--rate", type=float, default=20.0)
    ap.add_argument("--duration", type=float, default=120.0)
    ap.add_argument("--hmin", type=int, default=200)
    ap.add_argument("--hmax", type=int, default=2000)
    ap.add_argument("--budget", type=float, default=250e9)
    ap.add_argument("--check", type=float, default=0.05)
    a = ap.parse_args()
    if a.mode == "batch":
        out = batch(a.members, a.N, a.nlevs, a.steps)
    elif a.mode == "ensemble":
        out = ensemble(a.n, a.steps, a.budget, a.check)
    else:
        out = serve(a.rate, a.duration, a.hmin, a.hmax, a.budget, a.check)
    print(json.dumps(out))
    sys.stdout.flush()


if __name__ == "__main__":
    main()
