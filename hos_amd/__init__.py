"""Python binding of libhos_gpu.so (include/hos_gpu.h) for tests and the benchmark.

The product is the C-ABI library built from ``hos_amd/csrc`` (HIP, gfx950) and
the C host code in ``hos_amd/host``; this module only loads the library with
ctypes.  There is no CPU fallback: if the library is missing, or no GPU is
present, the calls fail loudly.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhos_gpu.so")
_lib = None

HOS_FP64 = 0
HOS_FP32 = 1
HOS_OK, HOS_E_INVALID, HOS_E_DEVICE, HOS_E_NOMEM, HOS_E_STATE = 0, -1, -2, -3, -4


class HosError(RuntimeError):
    pass


class Params(ctypes.Structure):
    _fields_ = [
        ("Nx", ctypes.c_int), ("Ny", ctypes.c_int), ("nlevs", ctypes.c_int),
        ("Lx", ctypes.c_double), ("Ly", ctypes.c_double), ("g", ctypes.c_double),
        ("depth", ctypes.c_double), ("rampflg", ctypes.c_int), ("Tramp", ctypes.c_double),
        ("scheme", ctypes.c_int), ("precision", ctypes.c_int), ("kx_quirk", ctypes.c_int),
        ("mean_level", ctypes.c_int), ("compensated", ctypes.c_int),
    ]


# every symbol include/hos_gpu.h declares
EXPORTS = [
    "hos_params_default", "hos_create", "hos_destroy", "hos_footprint", "hos_set_state", "hos_get_state",
    "hos_set_time", "hos_get_time", "hos_step", "hos_step_dt", "hos_get_energy", "hos_rhs", "hos_zvel",
    "hos_fft_2d", "hos_ifft_2d", "hos_sync", "hos_last_step_ms", "hos_use_graph", "hos_set_tile",
    "hos_last_error", "hos_device_name", "hos_num_members", "hos_bytes_per_point_step", "hos_step_times",
    "hos_profile", "hos_kernel_stats", "hos_spectral_op", "hos_mult", "hos_hamiltonian",
    "hos_set_active", "hos_get_active", "hos_set_member", "hos_copy_member", "hos_step_async", "hos_wait",
    "hos_autotune", "hos_copy_bandwidth", "hos_get_tile", "hos_query", "hos_rhs_noalloc",
]
(HOS_TILE_C2R, HOS_TILE_R2C, HOS_TILE_PAIR, HOS_TILE_ALL, HOS_TILE_XPASS, HOS_TILE_GROUPS, HOS_TILE_BAND,
 HOS_TILE_PLAN, HOS_TILE_FUSE, HOS_TILE_SKEW, HOS_TILE_DIAG, HOS_TILE_XBLOCK) = (0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11)
HOS_OP_DX, HOS_OP_DY, HOS_OP_DZ, HOS_OP_SUM, HOS_OP_DEALIAS = 1, 2, 3, 4, 5
# include/hos_ic.h
IC_EXPORTS = ["hos_ic_jonswap", "hos_ic_uniform", "hos_ic_peak_period", "hos_dealias_state"]

KERNEL_KINDS = ["xinv", "yc2r", "yr2c", "yr2c_pair", "xfwdinv", "xfwd_rhs", "yc2r_phi", "ydiag"]


class Jonswap(ctypes.Structure):
    _fields_ = [
        ("Nx", ctypes.c_int), ("Ny", ctypes.c_int), ("Lx", ctypes.c_double), ("Ly", ctypes.c_double),
        ("g", ctypes.c_double), ("depth", ctypes.c_double), ("kp", ctypes.c_double),
        ("gamma", ctypes.c_double), ("eps", ctypes.c_double), ("theta_p", ctypes.c_double),
        ("kcut", ctypes.c_double), ("seed", ctypes.c_ulonglong),
    ]


def build() -> str:
    """Compile libhos_gpu.so for gfx950 (hos_amd/Makefile)."""
    subprocess.run(["make", "-s", "-j8", "-C", _HERE, "-B", "all"], check=True)
    return LIB_PATH


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HosError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int
    P = ctypes.POINTER(Params)
    lib.hos_params_default.argtypes = [P]
    lib.hos_params_default.restype = None
    lib.hos_create.argtypes = [P, ip, ctypes.POINTER(vp)]
    lib.hos_destroy.argtypes = [vp]
    lib.hos_footprint.argtypes = [P, ip]
    lib.hos_footprint.restype = ctypes.c_size_t
    for n in ("hos_set_state", "hos_get_state"):
        getattr(lib, n).argtypes = [vp, ip, vp]
    lib.hos_set_time.argtypes = [vp, ip, ctypes.c_double]
    lib.hos_get_time.argtypes = [vp, ip, dp]
    lib.hos_step.argtypes = [vp, ip, ctypes.c_double]
    lib.hos_step_dt.argtypes = [vp, ip, vp]
    lib.hos_get_energy.argtypes = [vp, ip, dp, dp]
    lib.hos_rhs.argtypes = [vp, vp, vp, ctypes.c_double]
    lib.hos_zvel.argtypes = [vp, vp, vp, vp, vp, vp]
    lib.hos_fft_2d.argtypes = [vp, vp, vp]
    lib.hos_ifft_2d.argtypes = [vp, vp, vp]
    lib.hos_sync.argtypes = [vp]
    lib.hos_last_step_ms.argtypes = [vp, dp]
    lib.hos_use_graph.argtypes = [vp, ip]
    lib.hos_set_tile.argtypes = [vp, ip, ip]
    lib.hos_get_tile.argtypes = [vp, ip]
    lib.hos_last_error.restype = ctypes.c_char_p
    lib.hos_device_name.argtypes = [ctypes.c_char_p, ip]
    lib.hos_num_members.argtypes = [vp]
    lib.hos_bytes_per_point_step.argtypes = [P]
    lib.hos_bytes_per_point_step.restype = ctypes.c_double
    lib.hos_step_times.argtypes = [vp, vp, ip]
    lib.hos_profile.argtypes = [vp, ip]
    lib.hos_kernel_stats.argtypes = [vp, ip, dp, ctypes.POINTER(ctypes.c_long), dp]
    lib.hos_spectral_op.argtypes = [vp, ip, vp, vp, vp]
    lib.hos_mult.argtypes = [vp, vp, vp, vp]
    lib.hos_hamiltonian.argtypes = [vp, vp, vp, vp, dp]
    lib.hos_autotune.argtypes = [vp, ip, vp]
    lib.hos_set_active.argtypes = [vp, ip]
    lib.hos_get_active.argtypes = [vp]
    lib.hos_set_member.argtypes = [vp, ip, ctypes.c_double, ctypes.c_double, ctypes.c_double]
    lib.hos_copy_member.argtypes = [vp, ip, ip]
    lib.hos_step_async.argtypes = [vp, ip]
    lib.hos_wait.argtypes = [vp]
    lib.hos_query.argtypes = [vp]
    lib.hos_ic_jonswap.argtypes = [ctypes.POINTER(Jonswap), vp]
    lib.hos_ic_uniform.argtypes = [ctypes.c_ulonglong, ip, ip]
    lib.hos_ic_uniform.restype = ctypes.c_double
    lib.hos_ic_peak_period.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double]
    lib.hos_ic_peak_period.restype = ctypes.c_double
    lib.hos_dealias_state.argtypes = [ip, ip, ip, vp]
    lib.hos_dealias_state.restype = None
    _lib = lib
    return lib


def _check(rc: int):
    if rc != 0:
        raise HosError(f"hos error {rc}: {load().hos_last_error().decode()}")


def _ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return ctypes.c_void_p(a.ctypes.data)


def params(**kw) -> Params:
    p = Params()
    load().hos_params_default(ctypes.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def device_name() -> str:
    b = ctypes.create_string_buffer(256)
    _check(load().hos_device_name(b, 256))
    return b.value.decode()


def copy_bandwidth(nbytes: int = 1 << 30, reps: int = 20) -> float:
    """GB/s (read + write) of a plain device copy of nbytes (hos_copy_bandwidth)."""
    lib = load()
    lib.hos_copy_bandwidth.argtypes = [ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    v = ctypes.c_double()
    _check(lib.hos_copy_bandwidth(nbytes, reps, ctypes.byref(v)))
    return v.value


def footprint(p: Params, nmembers: int = 1) -> int:
    return int(load().hos_footprint(ctypes.byref(p), nmembers))


def bytes_per_point_step(p: Params) -> float:
    return float(load().hos_bytes_per_point_step(ctypes.byref(p)))


def jonswap(Nx, Ny, Lx, Ly, kp=1.0, gamma=3.3, eps=0.1, depth=0.0, g=9.81, theta_p=0.0, kcut=None,
            seed=0, nlevs=None) -> np.ndarray:
    """Seeded JONSWAP state [2][Ny/2+1][Nx] (hos_ic_jonswap); dealiased for
    nlevs when given (main_2d_par.c:258-262)."""
    lib = load()
    j = Jonswap(Nx, Ny, Lx, Ly, g, depth, kp, gamma, eps, theta_p, 4 * kp if kcut is None else kcut, seed)
    u = np.zeros((2, Ny // 2 + 1, Nx), np.complex128)
    if lib.hos_ic_jonswap(ctypes.byref(j), _ptr(u)) != 0:
        raise HosError("hos_ic_jonswap rejected its parameters")
    if nlevs is not None:
        lib.hos_dealias_state(Nx, Ny, nlevs, _ptr(u))
    return u


class ServerParams(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_double), ("precision", ctypes.c_int), ("use_graph", ctypes.c_int),
                ("max_slots", ctypes.c_int), ("compensated", ctypes.c_int), ("steps_per_launch", ctypes.c_int)]


class Request(ctypes.Structure):
    _fields_ = [("Nx", ctypes.c_int), ("Ny", ctypes.c_int), ("nlevs", ctypes.c_int), ("Lx", ctypes.c_double),
                ("Ly", ctypes.c_double), ("g", ctypes.c_double), ("depth", ctypes.c_double),
                ("dt", ctypes.c_double), ("nsteps", ctypes.c_int)]


SERVE_EXPORTS = ["hos_server_create", "hos_server_destroy", "hos_server_submit", "hos_server_status",
                 "hos_server_result", "hos_server_step", "hos_server_start", "hos_server_stop",
                 "hos_server_drain", "hos_server_stats", "hos_server_fail_launches"]
REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED = 0, 1, 2, 3


class Server:
    """hos_server (include/hos_serve.h): continuous batching of many runs."""

    def __init__(self, hbm_budget=250e9, precision=HOS_FP32, use_graph=True, max_slots=0, compensated=True,
                 steps_per_launch=0):
        lib = load()
        vp = ctypes.c_void_p
        lib.hos_server_create.argtypes = [ctypes.POINTER(ServerParams), ctypes.POINTER(vp)]
        lib.hos_server_destroy.argtypes = [vp]
        lib.hos_server_submit.argtypes = [vp, ctypes.POINTER(Request), vp, ctypes.POINTER(ctypes.c_longlong)]
        lib.hos_server_status.argtypes = [vp, ctypes.c_longlong]
        lib.hos_server_result.argtypes = [vp, ctypes.c_longlong, vp, vp]
        lib.hos_server_step.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
        lib.hos_server_start.argtypes = [vp]
        lib.hos_server_stop.argtypes = [vp]
        lib.hos_server_drain.argtypes = [vp, ctypes.c_double]
        lib.hos_server_stats.argtypes = [vp, vp, ctypes.c_int]
        self.lib = lib
        sp = ServerParams(float(hbm_budget), precision, int(use_graph), max_slots, int(compensated),
                          int(steps_per_launch))
        h = vp()
        if lib.hos_server_create(ctypes.byref(sp), ctypes.byref(h)) != 0:
            raise HosError("hos_server_create rejected its parameters")
        self.h = h
        self.shapes = {}

    def close(self):
        if getattr(self, "h", None):
            self.lib.hos_server_destroy(self.h)
            self.h = None

    __del__ = close

    def submit(self, u0, Nx, Ny, nlevs, Lx, Ly, dt, nsteps, depth=0.0, g=9.81) -> int:
        u0 = np.ascontiguousarray(u0, dtype=np.complex128)
        assert u0.shape == (2, Ny // 2 + 1, Nx)
        r = Request(Nx, Ny, nlevs, Lx, Ly, g, depth, dt, nsteps)
        rid = ctypes.c_longlong()
        rc = self.lib.hos_server_submit(self.h, ctypes.byref(r), _ptr(u0), ctypes.byref(rid))
        if rc != 0:
            raise HosError(f"submit rejected ({rc}): {load().hos_last_error().decode()}")
        self.shapes[rid.value] = (Nx, Ny)
        return rid.value

    def status(self, rid: int) -> int:
        return self.lib.hos_server_status(self.h, rid)

    def result(self, rid: int):
        Nx, Ny = self.shapes.pop(rid)
        u = np.zeros((2, Ny // 2 + 1, Nx), np.complex128)
        info = np.zeros(6)
        rc = self.lib.hos_server_result(self.h, rid, _ptr(u), _ptr(info))
        if rc != 0:
            raise HosError(f"request {rid} failed ({rc})")
        return u, dict(submit=info[0], admit=info[1], done=info[2], steps=int(info[3]), H=info[4], vol=info[5])

    def step(self) -> int:
        n = ctypes.c_int()
        _check(self.lib.hos_server_step(self.h, ctypes.byref(n)))
        return n.value

    def start(self):
        _check(self.lib.hos_server_start(self.h))

    def stop(self):
        _check(self.lib.hos_server_stop(self.h))

    def drain(self, timeout=600.0):
        _check(self.lib.hos_server_drain(self.h, timeout))

    def drain_for(self, seconds: float) -> bool:
        """hos_server_drain for at most `seconds`: True when nothing is queued
        or running any more, False when the time ran out first."""
        rc = self.lib.hos_server_drain(self.h, seconds)
        if rc == HOS_E_STATE:
            return False
        _check(rc)
        return True

    def fail_launches(self, n: int):
        """Failure injection for tests: the next n pool launches fail."""
        self.lib.hos_server_fail_launches.argtypes = [ctypes.c_void_p, ctypes.c_int]
        _check(self.lib.hos_server_fail_launches(self.h, n))

    def stats(self) -> dict:
        v = np.zeros(12)
        _check(self.lib.hos_server_stats(self.h, _ptr(v), 12))
        keys = ["step_p50_ms", "step_p99_ms", "req_p50_s", "req_p99_s", "completed", "pt_steps", "busy_s",
                "pools", "peak_bytes", "peak_running", "wait_p50_s", "wait_p99_s"]
        return dict(zip(keys, v.tolist()))


def peak_period(g, kp, depth=0.0) -> float:
    return load().hos_ic_peak_period(g, kp, depth)


class Solver:
    """A context of ``nmembers`` instances of one shape (hos_create)."""

    def __init__(self, p: Params, nmembers: int = 1):
        self.lib = load()
        self.p = p
        self.nm = nmembers
        self.Nx, self.Ny, self.Nk = p.Nx, p.Ny, p.Ny // 2 + 1
        h = ctypes.c_void_p()
        _check(self.lib.hos_create(ctypes.byref(p), nmembers, ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.hos_destroy(self.h)
            self.h = None

    __del__ = close

    def spec(self, n=1):
        shape = (self.Nk, self.Nx) if n == 1 else (n, self.Nk, self.Nx)
        return np.zeros(shape, dtype=np.complex128)

    def set_state(self, u: np.ndarray, member: int = 0):
        u = np.ascontiguousarray(u, dtype=np.complex128)
        assert u.shape == (2, self.Nk, self.Nx)
        _check(self.lib.hos_set_state(self.h, member, _ptr(u)))

    def get_state(self, member: int = 0) -> np.ndarray:
        u = self.spec(2)
        _check(self.lib.hos_get_state(self.h, member, _ptr(u)))
        return u

    def set_time(self, t: float, member: int = 0):
        _check(self.lib.hos_set_time(self.h, member, t))

    def get_time(self, member: int = 0) -> float:
        t = ctypes.c_double()
        _check(self.lib.hos_get_time(self.h, member, ctypes.byref(t)))
        return t.value

    def step(self, nsteps: int, dt):
        if np.isscalar(dt):
            _check(self.lib.hos_step(self.h, nsteps, float(dt)))
        else:
            d = np.ascontiguousarray(dt, dtype=np.float64)
            assert d.shape == (self.nm,)
            _check(self.lib.hos_step_dt(self.h, nsteps, _ptr(d)))

    def energy(self, member: int = 0):
        H, v = ctypes.c_double(), ctypes.c_double()
        _check(self.lib.hos_get_energy(self.h, member, ctypes.byref(H), ctypes.byref(v)))
        return H.value, v.value

    def rhs(self, u: np.ndarray, t: float = 0.0) -> np.ndarray:
        u = np.ascontiguousarray(u, dtype=np.complex128)
        out = self.spec(2)
        _check(self.lib.hos_rhs(self.h, _ptr(u), _ptr(out), t))
        return out

    def zvel(self, u: np.ndarray):
        u = np.ascontiguousarray(u, dtype=np.complex128)
        outs = [self.spec() for _ in range(4)]
        _check(self.lib.hos_zvel(self.h, _ptr(u), *[_ptr(o) for o in outs]))
        return tuple(outs)

    def fft_2d(self, u: np.ndarray) -> np.ndarray:
        up = np.zeros((self.Nx, self.Ny + 2))
        up[:, : self.Ny] = u
        hu = self.spec()
        _check(self.lib.hos_fft_2d(self.h, _ptr(up), _ptr(hu)))
        return hu

    def ifft_2d(self, hu: np.ndarray) -> np.ndarray:
        hu = np.ascontiguousarray(hu, dtype=np.complex128)
        up = np.zeros((self.Nx, self.Ny + 2))
        _check(self.lib.hos_ifft_2d(self.h, _ptr(hu), _ptr(up)))
        return up[:, : self.Ny].copy()

    def last_step_ms(self) -> float:
        ms = ctypes.c_double()
        _check(self.lib.hos_last_step_ms(self.h, ctypes.byref(ms)))
        return ms.value

    def set_active(self, n: int):
        _check(self.lib.hos_set_active(self.h, n))

    def set_member(self, m: int, depth: float, dt: float, t: float = 0.0):
        _check(self.lib.hos_set_member(self.h, m, depth, dt, t))

    def copy_member(self, src: int, dst: int):
        _check(self.lib.hos_copy_member(self.h, src, dst))

    def step_async(self, nsteps: int):
        _check(self.lib.hos_step_async(self.h, nsteps))

    def wait(self):
        _check(self.lib.hos_wait(self.h))

    def query(self) -> bool:
        """hos_query: True while steps of step_async are still running."""
        r = self.lib.hos_query(self.h)
        if r < 0:
            raise HosError(self.lib.hos_last_error().decode())
        return r == 1

    def spectral_op(self, op: int, a, b=None) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=np.complex128)
        bb = None if b is None else np.ascontiguousarray(b, dtype=np.complex128)
        out = self.spec()
        _check(self.lib.hos_spectral_op(self.h, op, _ptr(a), None if bb is None else _ptr(bb), _ptr(out)))
        return out

    def mult(self, a, b) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=np.complex128)
        b = np.ascontiguousarray(b, dtype=np.complex128)
        out = self.spec()
        _check(self.lib.hos_mult(self.h, _ptr(a), _ptr(b), _ptr(out)))
        return out

    def hamiltonian(self, heta, heta_t, hphi) -> float:
        a = [np.ascontiguousarray(x, dtype=np.complex128) for x in (heta, heta_t, hphi)]
        H = ctypes.c_double()
        _check(self.lib.hos_hamiltonian(self.h, *[_ptr(x) for x in a], ctypes.byref(H)))
        return H.value

    def step_times(self) -> np.ndarray:
        n = self.lib.hos_step_times(self.h, None, 0)
        out = np.zeros(max(n, 1))
        self.lib.hos_step_times(self.h, _ptr(out), n)
        return out[:n]

    def profile(self, on: bool = True):
        _check(self.lib.hos_profile(self.h, int(on)))

    def kernel_stats(self) -> dict:
        out = {}
        for k, name in enumerate(KERNEL_KINDS):
            ms, n, by = ctypes.c_double(), ctypes.c_long(), ctypes.c_double()
            _check(self.lib.hos_kernel_stats(self.h, k, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(by)))
            out[name] = {"ms": ms.value, "launches": n.value, "bytes": by.value}
        return out

    def use_graph(self, on: bool = True):
        _check(self.lib.hos_use_graph(self.h, int(on)))

    def autotune(self, nsteps: int = 2):
        """hos_autotune: returns the chosen (c2r, r2c, pair) y tiles; the
        x-pass mode it chose is get_tile(HOS_TILE_XPASS)."""
        t = (ctypes.c_int * 3)()
        _check(self.lib.hos_autotune(self.h, nsteps, t))
        return tuple(t)

    def set_tile(self, tx: int, kind: int = 3):
        """y-pass tile (4, 8, 16; 0 = default) for kind HOS_TILE_C2R (0),
        HOS_TILE_R2C (1), HOS_TILE_PAIR (2) or all (3); for HOS_TILE_XPASS (4)
        the x-pass mode (1 rows per workgroup, 2 row per wave, 0 default); for
        HOS_TILE_GROUPS (5) the member groups (1..4, 0 default)."""
        _check(self.lib.hos_set_tile(self.h, kind, tx))

    def get_tile(self, kind: int) -> int:
        v = self.lib.hos_get_tile(self.h, kind)
        if v < 0:
            raise HosError(self.lib.hos_last_error().decode())
        return v

    def sync(self):
        _check(self.lib.hos_sync(self.h))
