"""hos_rhs_noalloc (the allocation-free RHS) against hos_rhs on the same
inputs, through the C ABI with device buffers from the HIP runtime the library
itself uses: FP64 within 1e-13 relative L2 (the two differ only in how the FFT
twiddle table is computed: on the host in long double for hos_rhs, on the
device in FP64 here), FP32 within 1e-6.  Two calls in a row on the same
scratch array, the second with a smaller input (its stale contents from the
first call must not leak into the result), and a configuration above the
scratch limit is rejected with HOS_E_INVALID before any launch."""
import ctypes

import numpy as np
import pytest

from conftest import rel_l2
from helpers import random_state

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


@pytest.fixture(scope="module")
def hip(hos):
    lib = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so.7")  # the runtime libhos_gpu.so links against
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipFree.argtypes = [ctypes.c_void_p]
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    lib.hipDeviceSynchronize.argtypes = []
    hos.load().hos_rhs_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                           ctypes.c_void_p]
    return lib


def dev_buf(hip, nbytes):
    p = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(p), nbytes) == 0
    return p


def rhs_noalloc(hos, hip, p, u, t):
    n = u.nbytes
    du, dr = dev_buf(hip, n), dev_buf(hip, n)
    try:
        assert hip.hipMemcpy(du, u.ctypes.data, n, H2D) == 0
        rc = hos.load().hos_rhs_noalloc(ctypes.byref(p), du, dr, t, None)
        assert rc == 0, hos.load().hos_last_error().decode()
        assert hip.hipDeviceSynchronize() == 0
        out = np.zeros_like(u)
        assert hip.hipMemcpy(out.ctypes.data, dr, n, D2H) == 0
    finally:
        hip.hipFree(du)
        hip.hipFree(dr)
    return out


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_rhs_matches_hos_rhs(hos, hip, prec):
    tol = 1e-13 if prec == 0 else 1e-6
    cases = [  # (Nx, Ny, nlevs, depth, ramp, t): the second call is smaller
        (256, 128, 4, 0.9, 1, 1.3),
        (64, 32, 2, 0.0, 0, 0.0),
    ]
    for Nx, Ny, nlevs, depth, ramp, t in cases:
        print(f"noalloc prec={prec} {Nx}x{Ny} nlevs={nlevs}", flush=True)
        Lx, Ly = 2 * np.pi * Nx / 32, 2 * np.pi * Ny / 32
        p = hos.params(Nx=Nx, Ny=Ny, nlevs=nlevs, Lx=Lx, Ly=Ly, precision=prec, depth=depth, rampflg=ramp,
                       Tramp=2.0)
        u = np.ascontiguousarray(random_state(Nx, Ny, nlevs, amp=0.06, seed=Nx + nlevs, Lx=Lx, Ly=Ly))
        s = hos.Solver(p)
        ref = s.rhs(u, t)
        s.close()
        got = rhs_noalloc(hos, hip, p, u, t)
        err = rel_l2(got, ref)
        print(f"  rel L2 vs hos_rhs {err:.2e}", flush=True)
        assert err < tol, err


def test_noalloc_rejects_above_the_scratch_limit(hos, hip):
    p = hos.params(Nx=1024, Ny=1024, nlevs=4, precision=0)
    lib = hos.load()
    n = 2 * 513 * 1024 * 16
    du, dr = dev_buf(hip, n), dev_buf(hip, n)  # real buffers of the right size all the same
    try:
        rc = lib.hos_rhs_noalloc(ctypes.byref(p), du, dr, 0.0, None)
        assert hip.hipDeviceSynchronize() == 0
    finally:
        hip.hipFree(du)
        hip.hipFree(dr)
    assert rc == -1  # HOS_E_INVALID
    assert b"scratch" in lib.hos_last_error()
