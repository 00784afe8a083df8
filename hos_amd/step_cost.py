"""Per-step cost of the reference call boundary against native stepping.

What a driver written against Csource/2dpar pays per sol_update_RK call,
against the native C ABI on the same sea state: a hos_mkic JONSWAP state
(256^2 over 16 x 16 peak wavelengths, kp = 1, deep water, gamma = 3.3,
kp*Hs/2 = 0.1, seed 0), RK4, dt = Tp/64, NLevs = 2, FP64.

  python -m hos_amd.step_cost [--dir D]

- hos_driver (the reference driver's call sequence over libhos_ref.so: one
  synchronous sol_update_RK per step, the stage-0 H read back, a runtime.dat
  line) is run for 20 and for 220 steps, each twice; the difference of the
  best times over the 200 extra steps removes the context set-up, the
  initial Zvel and the HDF5 snapshot of t = 0.
- native: the same state through the C ABI, 200 calls of hos_step(c, 1)
  each followed by hos_get_energy, then one hos_step(c, 200), stream-ordered
  and with captured graphs.

Prints one JSON line (ms per RK4 step of each path).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import tempfile
import time

import hos_amd as hos

HERE = os.path.dirname(os.path.abspath(__file__))
N, NLEVS, KP, G = 256, 2, 1.0, 9.81


def driver_seconds(d, steps, dt):
    for f in os.listdir(d):
        if f.startswith("data") or f == "runtime.dat" or f.endswith(".xmf"):
            os.remove(os.path.join(d, f))
    t0 = time.perf_counter()
    subprocess.run([os.path.join(HERE, "hos_driver"), "-d", d, "-n", str(NLEVS), "-s", "2", "-t", repr(dt), "-N",
                    str(steps)], check=True, stdout=subprocess.DEVNULL)
    return time.perf_counter() - t0


def native(dt, u0, graph):
    L = 16 * 2 * math.pi / KP
    s = hos.Solver(hos.params(Nx=N, Ny=N, nlevs=NLEVS, Lx=L, Ly=L, g=G, precision=hos.HOS_FP64))
    s.use_graph(graph)
    s.set_state(u0)
    s.step(5, dt)
    s.sync()
    t0 = time.perf_counter()
    for _ in range(200):
        s.step(1, dt)
        s.energy(0)
    per_call = (time.perf_counter() - t0) / 200
    t0 = time.perf_counter()
    s.step(200, dt)
    s.sync()
    batched = (time.perf_counter() - t0) / 200
    s.close()
    return 1e3 * per_call, 1e3 * batched


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--dir", default=None, help="scratch folder for the driver's input and output")
    args = ap.parse_args()
    d = args.dir or tempfile.mkdtemp(prefix="hos_step_cost_")
    os.makedirs(d, exist_ok=True)
    L = 16 * 2 * math.pi / KP
    dt = hos.peak_period(G, KP) / 64
    # the driver's input: initpars.h5 / initdata.1.h5 (T and dtsave beyond the run: no snapshot but t = 0)
    subprocess.run([os.path.join(HERE, "hos_mkic"), str(N), str(N), repr(L), repr(L), repr(G), "0", repr(KP), "3.3",
                    "0.1", "0", "1000", "1000", "1", d], check=True, stdout=subprocess.DEVNULL)
    short = min(driver_seconds(d, 20, dt) for _ in range(2))
    long_ = min(driver_seconds(d, 220, dt) for _ in range(2))
    u0 = hos.jonswap(N, N, L, L, kp=KP, gamma=3.3, eps=0.1, seed=0, nlevs=NLEVS)
    calls_s, batch_s = native(dt, u0, False)
    calls_g, batch_g = native(dt, u0, True)
    print(json.dumps({
        "grid": N, "M": NLEVS + 1, "dt": dt,
        "driver_ms_per_step": 1e3 * (long_ - short) / 200,
        "driver_20_s": short, "driver_220_s": long_,
        "native_calls_of_1_step_ms": {"stream": calls_s, "graph": calls_g},
        "native_one_call_200_steps_ms": {"stream": batch_s, "graph": batch_g},
    }))


if __name__ == "__main__":
    main()
