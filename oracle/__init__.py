"""ctypes binding of the FP64 CPU restatement in ``oracle/hos_oracle.c``.

TEST INFRASTRUCTURE: the oracle is the parity checker for the HIP library and
the CPU baseline timed by ``bench.py``.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
it; the product path (``hos_amd``) never does.

Arrays follow the reference's layouts (Csource/2dpar):
  * spectral fields: complex128 ``[Ny/2+1][Nx]`` (index ``Nx*j + i``,
    ``FFT_TRANSPOSE 1``, euler_2d_par.h:22);
  * the state ``u``: ``[2][Ny/2+1][Nx]`` (eta-hat, then phi-hat at offset
    ``alloc_local``, main_2d_par.c:150-158);
  * real fields for fft_2d / ifft_2d: float64 ``[Nx][Ny+2]`` (padded rows,
    fft_routines_2d_par.c:24-34).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_c_p = ctypes.c_void_p
_d_p = ctypes.POINTER(ctypes.c_double)


def build() -> str:
    """Compile liboracle.so with gcc (oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "liboracle.so"], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    lib.or_create.restype = _c_p
    lib.or_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                              ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int]
    lib.or_destroy.argtypes = [_c_p]
    lib.or_fft_2d.argtypes = [_c_p, _c_p, _c_p]
    lib.or_ifft_2d.argtypes = [_c_p, _c_p, _c_p]
    lib.or_Dx.argtypes = [_c_p, _c_p, _c_p]
    lib.or_Dy.argtypes = [_c_p, _c_p, _c_p]
    lib.or_Dz.argtypes = [_c_p, _c_p, _c_p]
    lib.or_Mult.argtypes = [_c_p, _c_p, _c_p, _c_p]
    lib.or_Sum.argtypes = [_c_p, _c_p, _c_p, _c_p]
    lib.or_Dealias.argtypes = [_c_p, _c_p]
    lib.or_dealias_mx.argtypes = [ctypes.c_int, ctypes.c_int]
    lib.or_dealias_mx.restype = ctypes.c_int
    lib.or_Zvel.argtypes = [_c_p, _c_p, _c_p, _c_p, _c_p, _c_p, ctypes.c_double]
    lib.or_ZvelLinear.argtypes = [_c_p, _c_p, _c_p]
    lib.or_RampFun.argtypes = [_c_p, ctypes.c_double]
    lib.or_RampFun.restype = ctypes.c_double
    lib.or_rhs_hos.argtypes = [_c_p, _c_p, _c_p, ctypes.c_double]
    lib.or_Hamiltonian.argtypes = [_c_p, _c_p, _c_p, _c_p]
    lib.or_Hamiltonian.restype = ctypes.c_double
    lib.or_Setup_TimeScheme.argtypes = [_c_p, ctypes.c_int]
    lib.or_Setup_TimeScheme.restype = ctypes.c_int
    lib.or_sol_update_RK.argtypes = [_c_p, _c_p, _d_p, ctypes.c_double]
    lib.or_sol_update_RK.restype = ctypes.c_double
    lib.or_run.argtypes = [_c_p, _c_p, _d_p, ctypes.c_double, ctypes.c_int, _c_p]
    lib.or_get_tableau.argtypes = [_c_p, _c_p, _c_p, _c_p]
    lib.or_get_tableau.restype = ctypes.c_int
    lib.or_set_threads.argtypes = [ctypes.c_int]
    lib.or_set_threads.restype = ctypes.c_int
    _lib = lib
    return lib


def _ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"], "arrays passed to the oracle must be C-contiguous"
    return ctypes.c_void_p(a.ctypes.data)


def set_threads(n: int) -> int:
    """OpenMP threads of the oracle's transforms (n <= 0: the default);
    returns the count in use.  Results do not depend on it."""
    return _load().or_set_threads(n)


def dealias_mx(n: int, nlevs: int) -> int:
    """mx = floor(0.5*N/(1+0.5*NLevs)), operators_par.c:290."""
    return _load().or_dealias_mx(n, nlevs)


class Oracle:
    """One solver instance (the reference's globals, held in a context)."""

    def __init__(self, Nx: int, Ny: int, nlevs: int, Lx: float, Ly: float, g: float = 9.81,
                 rampflg: int = 0, Tramp: float = 1.0, depth: float = 0.0, kx_quirk: bool = True,
                 scheme: int = 2):
        self.lib = _load()
        self.Nx, self.Ny, self.nlevs = Nx, Ny, nlevs
        self.Lx, self.Ly, self.g = Lx, Ly, g
        self.Nk = Ny // 2 + 1
        self.ctx = self.lib.or_create(Nx, Ny, nlevs, Lx, Ly, g, rampflg, Tramp, depth, int(kx_quirk))
        if not self.ctx:
            raise ValueError("or_create rejected the configuration")
        if self.lib.or_Setup_TimeScheme(self.ctx, scheme) != 0:
            raise ValueError(f"unknown scheme {scheme}")

    def __del__(self):
        if getattr(self, "ctx", None):
            self.lib.or_destroy(self.ctx)
            self.ctx = None

    # -- layouts ----------------------------------------------------------
    def spec(self, n: int = 1) -> np.ndarray:
        shape = (self.Nk, self.Nx) if n == 1 else (n, self.Nk, self.Nx)
        return np.zeros(shape, dtype=np.complex128)

    def pad(self, u: np.ndarray) -> np.ndarray:
        """[Nx][Ny] -> padded [Nx][Ny+2]."""
        p = np.zeros((self.Nx, self.Ny + 2))
        p[:, : self.Ny] = u
        return p

    # -- L1 / L2 -----------------------------------------------------------
    def fft_2d(self, u: np.ndarray) -> np.ndarray:
        """u: [Nx][Ny] real -> [Ny/2+1][Nx] complex, normalised by 1/(Nx*Ny)."""
        up = self.pad(np.ascontiguousarray(u, dtype=np.float64))
        hu = self.spec()
        self.lib.or_fft_2d(self.ctx, _ptr(up), _ptr(hu))
        return hu

    def ifft_2d(self, hu: np.ndarray) -> np.ndarray:
        hu = np.ascontiguousarray(hu, dtype=np.complex128)
        up = np.zeros((self.Nx, self.Ny + 2))
        self.lib.or_ifft_2d(self.ctx, _ptr(hu), _ptr(up))
        return up[:, : self.Ny].copy()

    def _unary(self, fn, hu):
        hu = np.ascontiguousarray(hu, dtype=np.complex128)
        out = self.spec()
        fn(self.ctx, _ptr(hu), _ptr(out))
        return out

    def Dx(self, hu):
        return self._unary(self.lib.or_Dx, hu)

    def Dy(self, hu):
        return self._unary(self.lib.or_Dy, hu)

    def Dz(self, hu):
        return self._unary(self.lib.or_Dz, hu)

    def Mult(self, a, b):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        b = np.ascontiguousarray(b, dtype=np.complex128)
        out = self.spec()
        self.lib.or_Mult(self.ctx, _ptr(a), _ptr(b), _ptr(out))
        return out

    def Dealias(self, hu):
        out = np.array(hu, dtype=np.complex128, order="C", copy=True)
        self.lib.or_Dealias(self.ctx, _ptr(out))
        return out

    # -- L3 ----------------------------------------------------------------
    def Zvel(self, u: np.ndarray, t: float = 0.0):
        u = np.ascontiguousarray(u, dtype=np.complex128)
        outs = [self.spec() for _ in range(4)]
        self.lib.or_Zvel(self.ctx, _ptr(u), *[_ptr(o) for o in outs], t)
        return tuple(outs)  # W^(M), W^(M-2), (W^2)^(M), (W^2)^(M-2)

    def rhs_hos(self, u: np.ndarray, t: float = 0.0) -> np.ndarray:
        u = np.ascontiguousarray(u, dtype=np.complex128)
        out = self.spec(2)
        self.lib.or_rhs_hos(self.ctx, _ptr(out), _ptr(u), t)
        return out

    def Hamiltonian(self, heta, heta_t, hphi) -> float:
        a = [np.ascontiguousarray(x, dtype=np.complex128) for x in (heta, heta_t, hphi)]
        return self.lib.or_Hamiltonian(self.ctx, *[_ptr(x) for x in a])

    def RampFun(self, t: float) -> float:
        return self.lib.or_RampFun(self.ctx, t)

    # -- L4 ----------------------------------------------------------------
    def tableau(self):
        """(a[S], b[S][S], c[S]) of the scheme set by Setup_TimeScheme."""
        a, b, c = np.zeros(7), np.zeros((7, 7)), np.zeros(7)
        S = self.lib.or_get_tableau(self.ctx, _ptr(a), _ptr(b), _ptr(c))
        return a[:S], b[:S, :S], c[:S]

    def step(self, u: np.ndarray, t: float, dt: float):
        """One sol_update_RK call in place on u; returns (t_new, H_stage0)."""
        assert u.dtype == np.complex128 and u.flags["C_CONTIGUOUS"]
        tt = ctypes.c_double(t)
        H = self.lib.or_sol_update_RK(self.ctx, _ptr(u), ctypes.byref(tt), dt)
        return tt.value, H

    def run(self, u: np.ndarray, t: float, dt: float, nsteps: int):
        """nsteps sol_update_RK calls in place on u; returns (t_new, H per step)."""
        assert u.dtype == np.complex128 and u.flags["C_CONTIGUOUS"]
        tt = ctypes.c_double(t)
        H = np.zeros(max(nsteps, 1))
        self.lib.or_run(self.ctx, _ptr(u), ctypes.byref(tt), dt, nsteps, _ptr(H))
        return tt.value, H[:nsteps]
