"""Committed golden fixtures (tests/golden/*.npz, written by make_golden.py):
the IC generator and the oracle must keep reproducing them (CPU), and the
HIP path must match them (GPU)."""
import math
import os

import numpy as np
import pytest

from conftest import rel_l2

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    z = np.load(os.path.join(HERE, f"{name}.npz"), allow_pickle=False)
    return {k: z[k] for k in z.files}


def expand(band, N, nlevs):
    mx = int(math.floor(0.5 * N / (1 + 0.5 * nlevs)))
    cols = np.r_[0:mx, N - mx + 1:N]
    u = np.zeros((2, N // 2 + 1, N), complex)
    tmp = np.zeros((2, mx, N), complex)
    tmp[:, :, cols] = band
    u[:, :mx] = tmp
    return u


def ic_of(hos, f):
    N, nlevs = int(f["N"]), int(f["nlevs"])
    return hos.jonswap(N, N, float(f["L"]), float(f["L"]), kp=float(f["kp"]), gamma=float(f["gamma"]),
                       eps=float(f["eps"]), depth=float(f["depth"]), g=float(f["g"]), seed=int(f["seed"]),
                       nlevs=nlevs)


@pytest.mark.parametrize("name", ["small", "cfg1"])
def test_ic_generator_reproduces_fixture(hos, name):
    f = load(name)
    u = ic_of(hos, f)
    assert np.array_equal(u, expand(f["ic"], int(f["N"]), int(f["nlevs"])))


def test_oracle_reproduces_small_fixture(hos, oracle_mod):
    f = load("small")
    N, nlevs, L = int(f["N"]), int(f["nlevs"]), float(f["L"])
    o = oracle_mod.Oracle(N, N, nlevs, L, L, g=float(f["g"]), depth=float(f["depth"]))
    u = np.ascontiguousarray(expand(f["ic"], N, nlevs))
    t, H = o.run(u, 0.0, float(f["dt"]), int(f["steps"]))
    assert rel_l2(u, expand(f["final"], N, nlevs)) < 1e-14
    assert np.allclose(H, f["H"], rtol=1e-13, atol=0)


@pytest.mark.gpu
@pytest.mark.parametrize("name,tol64", [("small", 1e-12), ("cfg1", 1e-11)])
def test_gpu_fp64_matches_fixture(hos, name, tol64):
    """Plain FP64 GPU path vs the FP64 oracle after all steps: relative L2 of
    eta and phi_s each <= tol (1e-11 after 200 RK4 steps for cfg1)."""
    f = load(name)
    N, nlevs, L = int(f["N"]), int(f["nlevs"]), float(f["L"])
    print(f"golden {name} fp64", flush=True)
    p = hos.params(Nx=N, Ny=N, nlevs=nlevs, Lx=L, Ly=L, g=float(f["g"]), depth=float(f["depth"]),
                   precision=hos.HOS_FP64)
    s = hos.Solver(p)
    s.set_state(expand(f["ic"], N, nlevs))
    s.use_graph(True)
    s.step(int(f["steps"]), float(f["dt"]))
    got = s.get_state()
    ref = expand(f["final"], N, nlevs)
    e_eta, e_phi = rel_l2(got[0], ref[0]), rel_l2(got[1], ref[1])
    print(f"fp64 {name}: eta {e_eta:.2e} phi {e_phi:.2e}")
    assert e_eta <= tol64
    assert e_phi <= tol64
    H, vol = s.energy()
    assert H == pytest.approx(float(f["H"][-1]), rel=1e-11)


@pytest.mark.gpu
@pytest.mark.parametrize("name,tol32", [("small", 3e-7), ("cfg1", 5e-7)])
def test_gpu_fp32_error_vs_fixture(hos, name, tol32):
    """FP32 form (FP64 state) vs the FP64 oracle after all steps (measured
    2.7e-8 / 3.0e-8 for small, 6.2e-8 / 4.4e-8 for cfg1's 200 RK4 steps on
    eta / phi_s; the bounds leave a factor of about 8)."""
    f = load(name)
    N, nlevs, L = int(f["N"]), int(f["nlevs"]), float(f["L"])
    print(f"golden {name} fp32", flush=True)
    p = hos.params(Nx=N, Ny=N, nlevs=nlevs, Lx=L, Ly=L, g=float(f["g"]), depth=float(f["depth"]),
                   precision=hos.HOS_FP32, compensated=1)
    s = hos.Solver(p)
    s.set_state(expand(f["ic"], N, nlevs))
    s.step(int(f["steps"]), float(f["dt"]))
    got = s.get_state()
    ref = expand(f["final"], N, nlevs)
    e_eta, e_phi = rel_l2(got[0], ref[0]), rel_l2(got[1], ref[1])
    print(f"fp32 {name}: eta {e_eta:.2e} phi {e_phi:.2e}")
    assert e_eta <= tol32 and e_phi <= tol32
