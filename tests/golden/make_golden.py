#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ with the FP64 oracle.

Each fixture is an .npz (no pickles) holding the seeded JONSWAP initial state
(hos_ic_jonswap, host C) and the oracle's state after n RK4 steps, restricted
to the dealiasing band (the only non-zero modes), plus the stage-0
Hamiltonian of every step.

  cfg1   256x256, 16x16 peak wavelengths, kp = 1, deep water, M = 3,
         JONSWAP gamma = 3.3, kp*Hs/2 = 0.10, seed 0, dt = Tp/64, 200 steps
  small  64x64, 4x4 peak wavelengths, kp*h = 1.2, M = 4, gamma = 5,
         kp*Hs/2 = 0.12, seed 7, dt = Tp/32, 40 steps

Run from the repository root: python tests/golden/make_golden.py [name ...]
"""
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

CASES = {
    "cfg1": dict(N=256, nwl=16, depth=0.0, nlevs=2, gamma=3.3, eps=0.10, seed=0, tp_div=64, steps=200),
    "small": dict(N=64, nwl=4, depth=1.2, nlevs=3, gamma=5.0, eps=0.12, seed=7, tp_div=32, steps=40),
}


def band_index(N, nlevs):
    mx = int(math.floor(0.5 * N / (1 + 0.5 * nlevs)))
    cols = np.r_[0:mx, N - mx + 1:N]
    return mx, cols


def make(name):
    import hos_amd
    import oracle

    c = CASES[name]
    N, g, kp = c["N"], 9.81, 1.0
    L = c["nwl"] * 2 * math.pi / kp
    u0 = hos_amd.jonswap(N, N, L, L, kp=kp, gamma=c["gamma"], eps=c["eps"], depth=c["depth"], g=g,
                         seed=c["seed"], nlevs=c["nlevs"])
    Tp = hos_amd.peak_period(g, kp, c["depth"])
    dt = Tp / c["tp_div"]
    o = oracle.Oracle(N, N, c["nlevs"], L, L, g=g, depth=c["depth"], kx_quirk=True, scheme=2)
    u = np.ascontiguousarray(u0.copy())
    t0 = time.time()
    t, H = o.run(u, 0.0, dt, c["steps"])
    mx, cols = band_index(N, c["nlevs"])
    full = np.zeros_like(u)
    full[:, :mx][:, :, cols] = u[:, :mx][:, :, cols]
    assert np.array_equal(full, u), "state left the dealiasing band"
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), f"{name}.npz")
    np.savez_compressed(out, N=N, L=L, g=g, kp=kp, depth=c["depth"], nlevs=c["nlevs"], gamma=c["gamma"],
                        eps=c["eps"], seed=c["seed"], dt=dt, steps=c["steps"], t=t,
                        ic=u0[:, :mx][:, :, cols], final=u[:, :mx][:, :, cols], H=H)
    print(f"{name}: {c['steps']} steps in {time.time() - t0:.1f} s -> {out} ({os.path.getsize(out)} B)")


def expand(band, N, nlevs):
    """band [2][mx][2mx-1] -> full [2][N/2+1][N]."""
    mx, cols = band_index(N, nlevs)
    u = np.zeros((2, N // 2 + 1, N), complex)
    tmp = np.zeros((2, mx, N), complex)
    tmp[:, :, cols] = band
    u[:, :mx] = tmp
    return u


if __name__ == "__main__":
    for n in (sys.argv[1:] or list(CASES)):
        make(n)
