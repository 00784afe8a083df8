"""GPU parity of every entry point of the reference's call boundary against the
FP64 oracle: the native spectral operators (Dx, Dy, Dz, Sum, Dealias), Mult
and the Hamiltonian, then the same names called through libhos_ref.so
(include/hos_ref.h) by ctypes, the way a driver written against
Csource/2dpar calls them, and the reference's own input pair
(Csource/2dpar/initpars.h5 + initdata.1.h5, kept as data under tests/golden/)
run through hos_driver.

Tolerances: pointwise FP64 operators <= 1e-15 relative (max norm; the same
products in the same order, tanh may differ in the last bit); the dealiasing
mask bit-exact (index work); FP64 transforms and products <= 1e-12 relative
L2; FP32 Mult <= 1e-6; the Hamiltonian <= 1e-13 (a different summation
order).
"""
import ctypes
import math
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import rel_l2
from h5read import have_h5dump, read
from helpers import dealias_mask, random_state

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "hos_amd")
GOLD = os.path.join(ROOT, "tests", "golden", "ref2dpar")

Nx, Ny = 64, 32
Lx, Ly = 4 * np.pi, 2 * np.pi


def maxrel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


def solver(hos, nlevs, prec, depth):
    p = hos.params(Nx=Nx, Ny=Ny, nlevs=nlevs, Lx=Lx, Ly=Ly, precision=prec, depth=depth)
    return hos.Solver(p)


@pytest.mark.parametrize("depth", [0.0, 0.7])
@pytest.mark.parametrize("prec", [0, 1])
def test_spectral_ops_vs_oracle(hos, oracle_mod, depth, prec):
    """hos_spectral_op DX/DY/DZ/SUM/DEALIAS (operators_par.c:20-325) on FP64
    fields, in both context precisions (the operators are FP64 in both)."""
    nlevs = 2
    print(f"spectral ops depth={depth} prec={prec}", flush=True)
    o = oracle_mod.Oracle(Nx, Ny, nlevs, Lx, Ly, depth=depth)
    s = solver(hos, nlevs, prec, depth)
    rng = np.random.default_rng(5)
    a = rng.standard_normal((Ny // 2 + 1, Nx)) + 1j * rng.standard_normal((Ny // 2 + 1, Nx))
    b = rng.standard_normal((Ny // 2 + 1, Nx)) + 1j * rng.standard_normal((Ny // 2 + 1, Nx))
    assert maxrel(s.spectral_op(hos.HOS_OP_DX, a), o.Dx(a)) <= 1e-15
    assert maxrel(s.spectral_op(hos.HOS_OP_DY, a), o.Dy(a)) <= 1e-15
    assert maxrel(s.spectral_op(hos.HOS_OP_DZ, a), o.Dz(a)) <= 1e-15
    assert np.array_equal(s.spectral_op(hos.HOS_OP_SUM, a, b), a + b)
    got = s.spectral_op(hos.HOS_OP_DEALIAS, a)
    assert np.array_equal(got, o.Dealias(a))
    # the mask itself, from the index formula of operators_par.c:290-322
    m = dealias_mask(Nx, Ny, nlevs)
    assert np.array_equal(got == 0, m) and np.array_equal(got[~m], a[~m])
    s.close()


@pytest.mark.parametrize("depth", [0.0, 0.7])
@pytest.mark.parametrize("prec", [0, 1])
def test_mult_and_hamiltonian_vs_oracle(hos, oracle_mod, depth, prec):
    """hos_mult (operators_par.c:210-235) and hos_hamiltonian
    (hosm_2d_par.c:462-523)."""
    nlevs = 3
    print(f"mult/hamiltonian depth={depth} prec={prec}", flush=True)
    o = oracle_mod.Oracle(Nx, Ny, nlevs, Lx, Ly, depth=depth)
    s = solver(hos, nlevs, prec, depth)
    u = random_state(Nx, Ny, nlevs, amp=0.1, seed=13, Lx=Lx, Ly=Ly)
    w = random_state(Nx, Ny, nlevs, amp=0.1, seed=14, Lx=Lx, Ly=Ly)
    ref = o.Mult(u[0], u[1])
    err = rel_l2(s.mult(u[0], u[1]), ref)
    print(f"  Mult rel L2 {err:.2e}", flush=True)
    assert err < (1e-12 if prec == 0 else 1e-6), err
    # a square (a = b), as rhs_hos uses it (hosm_2d_par.c:167)
    assert rel_l2(s.mult(w[1], w[1]), o.Mult(w[1], w[1])) < (1e-12 if prec == 0 else 1e-6)
    H = s.hamiltonian(u[0], w[0], u[1])
    Href = o.Hamiltonian(u[0], w[0], u[1])
    assert abs(H - Href) <= 1e-13 * abs(Href), (H, Href)
    s.close()


# ------------------------------------------------------------------ libhos_ref.so
class RefLib:
    """libhos_ref.so with the reference's globals set through ctypes."""

    def __init__(self):
        self.lib = ctypes.CDLL(os.path.join(BIN, "libhos_ref.so"))
        L = self.lib
        vp, d = ctypes.c_void_p, ctypes.c_double
        for n in ("Dx", "Dy", "Dz", "Dealias", "ZvelLinear", "fft_2d", "ifft_2d", "Mult", "Sum", "rhs_hos",
                  "Zvel", "Hamiltonian", "RampFun"):
            getattr(L, n).restype = None
        L.rhs_hos.argtypes = [vp, vp, d]
        L.Zvel.argtypes = [vp, vp, vp, vp, vp, d]
        L.Hamiltonian.restype = d
        L.Hamiltonian.argtypes = [vp, vp, vp]
        L.RampFun.restype = d
        L.RampFun.argtypes = [d]
        L.fft_2d.argtypes = [vp, vp, vp]
        L.ifft_2d.argtypes = [vp, vp, vp]
        L.sol_update_RK.argtypes = [vp, ctypes.POINTER(d), d, vp]
        L.hos_local_size_2d_transposed.restype = ctypes.c_ssize_t
        L.hos_local_size_2d_transposed.argtypes = [ctypes.c_ssize_t] * 2 + [vp] * 4

    def setup(self, nlevs, depth, rampflg=0, Tramp=1.0, g=9.81, scheme=2):
        L = self.lib
        L.rhs_hos_clean()
        for name, ty, v in (("Nx", ctypes.c_int, Nx), ("Ny", ctypes.c_int, Ny), ("rampflg", ctypes.c_int, rampflg),
                            ("Lx", ctypes.c_double, Lx), ("Ly", ctypes.c_double, Ly), ("g", ctypes.c_double, g),
                            ("Tramp", ctypes.c_double, Tramp), ("hos_nlevs", ctypes.c_int, nlevs),
                            ("hos_depth", ctypes.c_double, depth), ("hos_runtime_log", ctypes.c_int, 0),
                            ("hos_precision", ctypes.c_int, 0), ("hos_kx_quirk", ctypes.c_int, 1)):
            ty.in_dll(L, name).value = v
        ctypes.c_double.in_dll(L, "Kx0").value = 2 * np.pi / Lx
        ctypes.c_double.in_dll(L, "Ky0").value = 2 * np.pi / Ly
        L.hos_local_size_2d_transposed(Nx, Ny // 2 + 1, None, None, None, None)
        L.Setup_TimeScheme(scheme)
        L.rhs_hos_setup()


def P(a):
    assert a.flags["C_CONTIGUOUS"]
    return ctypes.c_void_p(a.ctypes.data)


@pytest.fixture(scope="module")
def ref():
    r = RefLib()
    yield r
    r.lib.rhs_hos_clean()


@pytest.mark.parametrize("depth", [0.0, 0.7])
def test_reference_names_vs_oracle(ref, oracle_mod, depth):
    """Every numerical entry of hos_ref.h (time_schemes_par.h:20-26,
    hosm_2d_par.h:19-40, operators_par.h:17-23, fft_routines_2d_par.h:40-50)
    called by its reference name on host arrays, FP64, against the oracle's
    function of the same name."""
    nlevs = 2
    print(f"reference names depth={depth}", flush=True)
    ref.setup(nlevs, depth, rampflg=1, Tramp=2.0)
    L = ref.lib
    o = oracle_mod.Oracle(Nx, Ny, nlevs, Lx, Ly, depth=depth, rampflg=1, Tramp=2.0)
    u = random_state(Nx, Ny, nlevs, amp=0.08, seed=31, Lx=Lx, Ly=Ly)
    spec = lambda: np.zeros((Ny // 2 + 1, Nx), complex)  # noqa: E731

    def un(name, a):
        out = spec()
        getattr(L, name)(P(np.ascontiguousarray(a)), P(out))
        return out

    assert maxrel(un("Dx", u[0]), o.Dx(u[0])) <= 1e-15
    assert maxrel(un("Dy", u[1]), o.Dy(u[1])) <= 1e-15
    assert maxrel(un("Dz", u[1]), o.Dz(u[1])) <= 1e-15
    zl = spec()
    L.ZvelLinear(P(u), P(zl))  # Dz of phi-hat (hosm_2d_par.c:448-455)
    ozl = np.zeros_like(zl)
    o.lib.or_ZvelLinear(o.ctx, P(u), P(ozl))
    assert maxrel(zl, ozl) <= 1e-15
    s = spec()
    a0, a1 = np.ascontiguousarray(u[0]), np.ascontiguousarray(u[1])
    L.Sum(P(a0), P(a1), P(s))
    assert np.array_equal(s, a0 + a1)
    rng = np.random.default_rng(2)
    full = np.ascontiguousarray(rng.standard_normal(s.shape) + 1j * rng.standard_normal(s.shape))
    want = o.Dealias(full)
    L.Dealias(P(full))
    assert np.array_equal(full, want)
    # Mult in place (out aliases the first input), as hosm_2d_par.c:167
    prod = a0.copy()
    L.Mult(P(prod), P(a1), P(prod))
    assert rel_l2(prod, o.Mult(a0, a1)) < 1e-12
    # fft_2d / ifft_2d on the padded real layout [Nx][Ny+2]
    x = np.random.default_rng(3).standard_normal((Nx, Ny))
    xp = np.zeros((Nx, Ny + 2))
    xp[:, :Ny] = x
    hx = spec()
    L.fft_2d(P(xp), P(hx), None)
    assert rel_l2(hx, o.fft_2d(x)) < 1e-14
    back = np.zeros((Nx, Ny + 2))
    L.ifft_2d(P(hx), P(back), None)
    assert rel_l2(back[:, :Ny], o.ifft_2d(hx)) < 1e-14
    # rhs_hos with the ramp, Zvel, Hamiltonian, RampFun
    for t in (0.0, 1.3):
        r = np.zeros((2, Ny // 2 + 1, Nx), complex)
        L.rhs_hos(P(r), P(u), t)
        assert rel_l2(r, o.rhs_hos(u, t)) < 1e-12, t
        assert L.RampFun(t) == o.RampFun(t)
    outs = [spec() for _ in range(4)]
    L.Zvel(P(u), *[P(w) for w in outs], 0.0)
    for g_, r_ in zip(outs, o.Zvel(u)):
        if np.linalg.norm(r_) == 0:  # (W^2)^(M-2) at M = 3 has no terms
            assert np.linalg.norm(g_) < 1e-14
        else:
            assert rel_l2(g_, r_) < 1e-12
    et = np.ascontiguousarray(outs[0])
    H = L.Hamiltonian(P(a0), P(et), P(a1))
    Href = o.Hamiltonian(a0, et, a1)
    assert abs(H - Href) <= 1e-13 * abs(Href)


def test_reference_sol_update_rk_resident_state(ref, oracle_mod):
    """sol_update_RK keeps the state on the GPU between calls; a boundary call
    that reads the host state (ifft_2d of it, Zvel of it) sees the latest step,
    and a Zvel writing into the state array marks the GPU copy stale."""
    nlevs = 2
    print("reference sol_update_RK residency", flush=True)
    ref.setup(nlevs, 0.0, scheme=2)
    L = ref.lib
    o = oracle_mod.Oracle(Nx, Ny, nlevs, Lx, Ly)
    u = random_state(Nx, Ny, nlevs, amp=0.06, seed=33, Lx=Lx, Ly=Ly)
    ur = u.copy()
    t = ctypes.c_double(0.0)
    tr = 0.0
    for _ in range(3):
        L.sol_update_RK(P(u), ctypes.byref(t), 0.05, None)
        tr, _ = o.step(ur, tr, 0.05)
    assert t.value == pytest.approx(tr)
    eta = np.zeros((Nx, Ny + 2))
    L.ifft_2d(P(u), P(eta), None)  # reads heta = &u[0]: syncs the host copy first
    assert rel_l2(u, ur) < 1e-12
    assert rel_l2(eta[:, :Ny], o.ifft_2d(ur[0])) < 1e-12
    # Zvel writing its W^(M) output over phi-hat, inside the registered state:
    # the next step starts from it without any call from the caller
    scratch = [np.zeros((Ny // 2 + 1, Nx), complex) for _ in range(3)]
    ur[1] = o.Zvel(ur)[0]
    L.Zvel(P(u), P(u[1]), *[P(s) for s in scratch], 0.0)
    assert rel_l2(u[1], ur[1]) < 1e-12
    L.sol_update_RK(P(u), ctypes.byref(t), 0.05, None)
    o.step(ur, tr, 0.05)
    L.hos_ref_sync()
    assert rel_l2(u, ur) < 1e-12


# --------------------------------------------------------- reference input pair
@pytest.mark.skipif(not have_h5dump(), reason="h5dump not in this image")
def test_reference_input_pair_through_driver(tmp_path, oracle_mod):
    """The reference's own initpars.h5 / initdata.1.h5 (256^2, Lx = 3807.99,
    rampflg = 1, Tramp = 200) through hos_driver with the reference driver's
    own choices (midpoint scheme 4, dt = 0.2, NLevs = 2; main_2d_par.c:252,302,
    euler_2d_par.h:24) for 20 steps, with the extra W output on (-S 2): every
    snapshot, the t = 0 W^(M), W^(M-2) files and runtime.dat against the
    oracle on the same input."""
    print("reference input pair through hos_driver", flush=True)
    d = str(tmp_path)
    for f in ("initpars.h5", "initdata.1.h5"):
        shutil.copy(os.path.join(GOLD, f), d)
    nsteps, dt = 20, 0.2
    out = subprocess.run([os.path.join(BIN, "hos_driver"), "-d", d, "-N", str(nsteps), "-S", "2"],
                         check=True, capture_output=True, text=True, timeout=120)
    print(out.stdout[-300:], flush=True)
    n = 256
    Lx_ = read(os.path.join(d, "data0.1.h5"), "/Lx")[0]
    assert Lx_ == pytest.approx(3807.991095260355)
    eta0 = read(os.path.join(d, "initdata.1.h5"), "/eta0", (n, n))
    phi0 = read(os.path.join(d, "initdata.1.h5"), "/phi0", (n, n))
    o = oracle_mod.Oracle(n, n, 2, Lx_, Lx_, g=9.81, rampflg=1, Tramp=200.0, scheme=4)
    u = np.zeros((2, n // 2 + 1, n), complex)
    u[0] = o.Dealias(o.fft_2d(eta0))
    u[1] = o.Dealias(o.fft_2d(phi0))
    # data_extra0: physical W^(M), W^(M-2) of the dealiased IC (main_2d_par.c:264-294)
    wM, wM2, _, _ = o.Zvel(u)
    x0 = os.path.join(d, "data_extra0.1.h5")
    assert rel_l2(read(x0, "/Array1", (n, n)), o.ifft_2d(wM)) < 1e-12
    assert rel_l2(read(x0, "/Array2", (n, n)), o.ifft_2d(wM2)) < 1e-12
    snaps = sorted(int(m.group(1)) for m in (re.fullmatch(r"data(\d+)\.1\.h5", f) for f in os.listdir(d)) if m)
    per = int(round(1.0 / dt))  # dtsave = 1
    assert snaps == list(range(nsteps // per + 1))
    t = 0.0
    Hs = []
    for k in range(1, nsteps // per + 1):
        t, H = o.run(u, t, dt, per)
        Hs.extend(H)
        f = os.path.join(d, f"data{k}.1.h5")
        assert read(f, "/time")[0] == pytest.approx(k * per * dt, rel=1e-12)
        assert rel_l2(read(f, "/eta", (n, n)), o.ifft_2d(u[0])) < 1e-12, k
        assert rel_l2(read(f, "/phi", (n, n)), o.ifft_2d(u[1])) < 1e-12, k
        # the stale extras of later snapshots (the reference never refreshes them)
        xk = os.path.join(d, f"data_extra{k}.1.h5")
        assert np.array_equal(read(xk, "/Array1", (n, n)), read(x0, "/Array1", (n, n)))
    rt = np.loadtxt(os.path.join(d, "runtime.dat"))
    assert rt.shape == (nsteps, 2)
    assert np.allclose(rt[:, 1], Hs, rtol=1e-12)
    # the raw IC in data0 (main_2d_par.c:271-273)
    assert np.array_equal(read(os.path.join(d, "data0.1.h5"), "/eta", (n, n)), eta0)
    assert math.isfinite(float(rt[-1, 1]))


def test_float_boundary_vs_oracle(oracle_mod):
    """libhos_ref_f32.so: the same names over float / fftwf_complex (the
    reference built without USE_DOUBLES, hosm_2d_par.h:31-38,
    operators_par.h:24-30, fft_routines_2d_par.h:45-47,
    time_schemes_par.h:22-24) on float host arrays, computing in the FP32
    form with the GPU state in FP64 between sol_update_RK calls; against the
    oracle in FP64 on the same (float-rounded) input.  Tolerance 2e-6
    relative L2: the float rounding of inputs and outputs (6e-8) plus the
    FP32 form (below 3e-7)."""
    print("float boundary", flush=True)
    nlevs = 2
    L = ctypes.CDLL(os.path.join(BIN, "libhos_ref_f32.so"))
    f, vp = ctypes.c_float, ctypes.c_void_p
    L.rhs_hos.argtypes = [vp, vp, f]
    L.rhs_hos.restype = None
    L.sol_update_RK.argtypes = [vp, ctypes.POINTER(f), f, vp]
    L.sol_update_RK.restype = None
    L.Hamiltonian.argtypes = [vp, vp, vp]
    L.Hamiltonian.restype = f
    L.hos_local_size_2d_transposed.restype = ctypes.c_ssize_t
    L.hos_local_size_2d_transposed.argtypes = [ctypes.c_ssize_t] * 2 + [vp] * 4
    for name, ty, v in (("Nx", ctypes.c_int, Nx), ("Ny", ctypes.c_int, Ny), ("rampflg", ctypes.c_int, 0),
                        ("Lx", f, Lx), ("Ly", f, Ly), ("g", f, 9.81), ("Tramp", f, 1.0),
                        ("hos_nlevs", ctypes.c_int, nlevs), ("hos_runtime_log", ctypes.c_int, 0)):
        ty.in_dll(L, name).value = v
    assert ctypes.c_int.in_dll(L, "hos_precision").value == 1  # HOS_FP32 by default in the float build
    L.hos_local_size_2d_transposed(Nx, Ny // 2 + 1, None, None, None, None)
    L.Setup_TimeScheme(2)
    L.rhs_hos_setup()
    Lx32, Ly32, g32 = (float(np.float32(v)) for v in (Lx, Ly, 9.81))
    o = oracle_mod.Oracle(Nx, Ny, nlevs, Lx32, Ly32, g=g32)
    u64 = random_state(Nx, Ny, nlevs, amp=0.08, seed=35, Lx=Lx, Ly=Ly)
    u = np.ascontiguousarray(u64.astype(np.complex64))  # fftwf_complex [2][Ny/2+1][Nx]
    ur = np.ascontiguousarray(u.astype(np.complex128))
    r = np.zeros_like(u)
    L.rhs_hos(P(r), P(u), 0.0)
    err = rel_l2(r, o.rhs_hos(ur))
    print(f"  rhs_hos rel L2 {err:.2e}", flush=True)
    assert err < 2e-6
    d = np.zeros_like(u[0])
    a = np.ascontiguousarray(u[1])
    L.Dx(P(a), P(d))
    assert rel_l2(d, o.Dx(np.ascontiguousarray(ur[1]))) < 2e-7
    t = f(0.0)
    tr = 0.0
    for _ in range(4):
        L.sol_update_RK(P(u), ctypes.byref(t), f(0.02), None)
        tr, H = o.step(ur, tr, float(np.float32(0.02)))
    L.hos_ref_sync()
    err = rel_l2(u, ur)
    print(f"  4 RK4 steps rel L2 {err:.2e}", flush=True)
    assert err < 2e-6
    assert abs(t.value - tr) < 1e-6
    L.rhs_hos_clean()
