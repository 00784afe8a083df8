"""Long-horizon accuracy of the FP32 form at the 1024^2, M=5, kp*h=1 workload:
after 1000 RK4 steps (dt = Tp/100) eta must agree with the plain FP64 path to
<= 1e-5 relative L2.  The FP64 path itself matches the FP64 oracle to
<= 1e-11 after 200 steps (test_golden.py); here it is additionally checked
against the oracle for one step at this size.  Energy (H) and volume drift
of both forms are printed and bounded."""
import math

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def workload(hos):
    kp = 1.0
    L = 32 * 2 * math.pi / kp
    u0 = hos.jonswap(1024, 1024, L, L, kp=kp, gamma=5.0, eps=0.12, depth=1.0, seed=1, nlevs=4)
    dt = hos.peak_period(9.81, kp, 1.0) / 100
    return L, u0, dt


def run(hos, prec, u0, L, dt, nsteps, chunk=100):
    p = hos.params(Nx=1024, Ny=1024, nlevs=4, Lx=L, Ly=L, depth=1.0, precision=prec, compensated=1)
    s = hos.Solver(p)
    s.set_state(u0)
    s.use_graph(True)
    Hs, vols = [], []
    for _ in range(nsteps // chunk):
        s.step(chunk, dt)
        H, v = s.energy()
        Hs.append(H)
        vols.append(v)
    u = s.get_state()
    s.close()
    return u, np.array(Hs), np.array(vols)


def test_fp32_1000_steps_vs_fp64(hos):
    L, u0, dt = workload(hos)
    print("accuracy 1000 steps fp32", flush=True)
    u32, H32, v32 = run(hos, hos.HOS_FP32, u0, L, dt, 1000)
    print("accuracy 1000 steps fp64", flush=True)
    u64, H64, v64 = run(hos, hos.HOS_FP64, u0, L, dt, 1000)
    e_eta = rel_l2(u32[0], u64[0])
    e_phi = rel_l2(u32[1], u64[1])
    dH32 = abs(H32[-1] - H64[0]) / abs(H64[0])
    dH64 = abs(H64[-1] - H64[0]) / abs(H64[0])
    print(f"1000 steps: eta rel L2 {e_eta:.3e}, phi {e_phi:.3e}; H drift fp32 {dH32:.3e} fp64 {dH64:.3e}; "
          f"volume fp32 {v32[-1]:.3e} fp64 {v64[-1]:.3e}")
    assert e_eta <= 1e-5 and e_phi <= 1e-5  # the stated accuracy target
    # regression guard: measured 2.5e-7 (eta) and 3.1e-7 (phi_s)
    assert e_eta <= 2e-6 and e_phi <= 2e-6
    assert abs(v32[-1] - v64[-1]) < 1e-9
    assert dH64 < 1e-3 and dH32 < 1e-3


def test_fp64_one_step_vs_oracle_1024(hos, oracle_mod):
    L, u0, dt = workload(hos)
    print("fp64 1024^2 step vs oracle", flush=True)
    p = hos.params(Nx=1024, Ny=1024, nlevs=4, Lx=L, Ly=L, depth=1.0, precision=hos.HOS_FP64)
    s = hos.Solver(p)
    s.set_state(u0)
    s.step(1, dt)
    o = oracle_mod.Oracle(1024, 1024, 4, L, L, depth=1.0)
    ref = np.ascontiguousarray(u0.copy())
    o.run(ref, 0.0, dt, 1)
    err = rel_l2(s.get_state(), ref)
    print(f"fp64 1024^2 one RK4 step vs oracle: {err:.2e}")
    assert err < 1e-13


def test_fp32_batch_1000_steps_vs_fp64(hos):
    """The benchmark's form: eight distinct sea states (seeds 1-8) in one FP32
    context with one row per wave in the x passes, blocked x-intermediates and
    two member groups under graphs.  After 1000 RK4 steps members 0 and 7 agree
    with the plain FP64 path of the same states to <= 1e-5 (eta and phi_s)."""
    kp = 1.0
    L = 32 * 2 * math.pi / kp
    dt = hos.peak_period(9.81, kp, 1.0) / 100
    nm = 8
    states = [hos.jonswap(1024, 1024, L, L, kp=kp, gamma=5.0, eps=0.12, depth=1.0, seed=1 + m, nlevs=4)
              for m in range(nm)]
    print("accuracy 1000 steps fp32 batch of 8", flush=True)
    p = hos.params(Nx=1024, Ny=1024, nlevs=4, Lx=L, Ly=L, depth=1.0, precision=hos.HOS_FP32, compensated=1)
    s = hos.Solver(p, nm)
    for m, u in enumerate(states):
        s.set_state(u, m)
    s.set_tile(2, hos.HOS_TILE_XPASS)
    s.set_tile(2, hos.HOS_TILE_GROUPS)
    s.set_tile(2, hos.HOS_TILE_XBLOCK)
    s.use_graph(True)
    for _ in range(10):
        s.step(100, dt)
    got = {m: s.get_state(m) for m in (0, nm - 1)}
    s.close()
    for m in (0, nm - 1):
        print(f"  fp64 member {m}", flush=True)
        u64 = run(hos, hos.HOS_FP64, states[m], L, dt, 1000)[0]
        e_eta, e_phi = rel_l2(got[m][0], u64[0]), rel_l2(got[m][1], u64[1])
        print(f"  member {m}: eta rel L2 {e_eta:.3e}, phi {e_phi:.3e}", flush=True)
        assert e_eta <= 1e-5 and e_phi <= 1e-5
        assert e_eta <= 2e-6 and e_phi <= 2e-6  # regression guard (the single-member form: 2.5e-7)
