"""GPU parity of the diagonal passes (HOS_TILE_DIAG): the W expansion as
passes over the diagonals order + derivative = L+1, whose y c2r happens in the
pass that consumes them, and the RHS pass over W^(1) .. W^(M-1).

Against the FP64 oracle on the same inputs: FP64 <= 1e-12 relative L2 per RHS,
FP32 <= 2e-6 (the tolerances of the other RHS tests).  Against the mode off:
RK steps agree to rounding (FP64 <= 1e-13 after 10 steps), and members of a
batch are bit-identical to the same members alone with the mode on.
"""
import numpy as np
import pytest

from helpers import random_state
from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 2e-6


def solver(hos, Nx, Ny, nlevs, prec, nm=1, diag=2, **kw):
    p = hos.params(Nx=Nx, Ny=Ny, nlevs=nlevs, Lx=kw.pop("Lx", 2 * np.pi), Ly=kw.pop("Ly", 2 * np.pi),
                   precision=prec, **kw)
    s = hos.Solver(p, nm)
    s.set_tile(diag, hos.HOS_TILE_DIAG)
    assert s.get_tile(hos.HOS_TILE_DIAG) == diag
    return s


@pytest.mark.parametrize("nlevs", [0, 1, 2, 3, 4, 5, 7])
@pytest.mark.parametrize("depth", [0.0, 0.7])
def test_diag_rhs_fp64_vs_oracle(hos, oracle_mod, nlevs, depth):
    Nx, Ny = 64, 32
    Lx, Ly = 4 * np.pi, 2 * np.pi
    print(f"diag rhs fp64 nlevs={nlevs} depth={depth}", flush=True)
    o = oracle_mod.Oracle(Nx, Ny, nlevs, Lx, Ly, depth=depth)
    s = solver(hos, Nx, Ny, nlevs, 0, Lx=Lx, Ly=Ly, depth=depth)
    u = random_state(Nx, Ny, nlevs, amp=0.08, seed=nlevs)
    err = rel_l2(s.rhs(u), o.rhs_hos(u))
    print(f"  rel L2 {err:.2e}", flush=True)
    assert err < TOL64, err
    s.close()


@pytest.mark.parametrize("nlevs", [1, 2, 4, 5, 7])
@pytest.mark.parametrize("depth", [0.0, 0.9])
def test_diag_rhs_fp32_vs_oracle(hos, oracle_mod, nlevs, depth):
    """FP32 form, M = 2..8: up to three diagonal fields per LDS batch, one to
    three batches per pass; the RHS pass reads its W^(a) fields before the
    deferred ones at every M."""
    Nx, Ny = 128, 64
    print(f"diag rhs fp32 nlevs={nlevs} depth={depth}", flush=True)
    o = oracle_mod.Oracle(Nx, Ny, nlevs, 2 * np.pi, np.pi, depth=depth)
    s = solver(hos, Nx, Ny, nlevs, 1, Ly=np.pi, depth=depth)
    u = random_state(Nx, Ny, nlevs, amp=0.05, seed=7)
    err = rel_l2(s.rhs(u), o.rhs_hos(u))
    print(f"  rel L2 {err:.2e}", flush=True)
    assert err < TOL32, err
    s.close()


@pytest.mark.parametrize("Nx,Ny", [(2048, 32), (32, 2048), (256, 1024)])
@pytest.mark.parametrize("prec", [0, 1])
def test_diag_rhs_long_rows_vs_oracle(hos, oracle_mod, Nx, Ny, prec):
    """Rows of 2048 points (the FP64 tiles hold more than 64 KiB of LDS) and
    1024-point y rows, finite depth, ramp on."""
    nlevs, depth = 3, 0.9
    Lx, Ly = 2 * np.pi * Nx / 32, 2 * np.pi * Ny / 32
    print(f"diag rhs {Nx}x{Ny} prec={prec}", flush=True)
    o = oracle_mod.Oracle(Nx, Ny, nlevs, Lx, Ly, depth=depth, rampflg=1, Tramp=2.0)
    s = solver(hos, Nx, Ny, nlevs, prec, Lx=Lx, Ly=Ly, depth=depth, rampflg=1, Tramp=2.0)
    u = random_state(Nx, Ny, nlevs, amp=0.05, slope=3.0, seed=13)
    err = rel_l2(s.rhs(u, 1.3), o.rhs_hos(u, 1.3))
    print(f"  rel L2 {err:.2e}", flush=True)
    assert err < (TOL64 if prec == 0 else TOL32), err
    s.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_diag_rhs_fft_plans_vs_oracle(hos, oracle_mod, prec):
    """Every radix plan of the RHS pass over W^(a) fields (Ny >= 512)."""
    Nx, Ny, nlevs = 64, 512, 4
    print(f"diag rhs plans prec={prec}", flush=True)
    o = oracle_mod.Oracle(Nx, Ny, nlevs, 2 * np.pi, 16 * np.pi)
    s = solver(hos, Nx, Ny, nlevs, prec, Ly=16 * np.pi)
    u = random_state(Nx, Ny, nlevs, amp=0.05, slope=3.0, seed=17, Ly=16 * np.pi)
    ref = o.rhs_hos(u)
    for plan in (1, 2, 3, 4):
        s.set_tile(plan, hos.HOS_TILE_PLAN)
        err = rel_l2(s.rhs(u), ref)
        print(f"  plan {plan}: rel L2 {err:.2e}", flush=True)
        assert err < (TOL64 if prec == 0 else TOL32), (plan, err)
    s.close()


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("band", [1, 2])
@pytest.mark.parametrize("nlevs", [0, 2, 4])
def test_diag_steps_match_mode_off(hos, prec, band, nlevs):
    """Ten RK4 steps with the diagonal passes against the same steps with the
    mode off, a masked state with the band of masked states on and off, M = 1
    (no diagonal pass: eta's own y c2r), 3 and 5.  FP64: the two forms agree
    to rounding, states and energies.  FP32: each form's distance from the
    FP64 steps is the FP32 form's error (rounding amplified by the dynamics:
    at M = 1 the mask keeps nearly every mode); the two forms agree within a
    few times it."""
    Nx, Ny = 128, 128
    print(f"diag steps prec={prec} band={band} nlevs={nlevs}", flush=True)
    u = random_state(Nx, Ny, nlevs, amp=0.06, seed=23)

    def run(diag, pr):
        s = solver(hos, Nx, Ny, nlevs, pr, diag=diag, depth=1.2)
        s.set_tile(band, hos.HOS_TILE_BAND)
        s.set_state(u)
        s.step(10, 0.03)
        out = (s.get_state(), s.energy(0)[0])
        s.close()
        return out

    on, off = run(2, prec), run(1, prec)
    err = rel_l2(on[0], off[0])
    if prec == 0:
        print(f"  rel L2 {err:.2e}", flush=True)
        assert err < 1e-13, err
        assert abs(on[1] - off[1]) <= 1e-12 * abs(off[1])
    else:
        ref = run(2, 0)[0]
        e_on, e_off = rel_l2(on[0], ref), rel_l2(off[0], ref)
        print(f"  forms {err:.2e}; vs FP64: diagonal {e_on:.2e}, level {e_off:.2e}", flush=True)
        assert e_on < max(1e-6, 3 * e_off), (e_on, e_off)
        assert err < max(1e-6, 3 * e_off), (err, e_off)


@pytest.mark.parametrize("graph", [False, True])
def test_diag_members_match_alone(hos, graph):
    """With the diagonal passes, members of a batch over two offset member
    groups are bit-identical to the same members stepped alone."""
    Nx, Ny, nlevs, nm = 64, 64, 4, 5
    print(f"diag members graph={graph}", flush=True)
    states = [random_state(Nx, Ny, nlevs, amp=0.06, seed=70 + i) for i in range(nm)]
    dts = 0.02 + 0.001 * np.arange(nm)
    batch = solver(hos, Nx, Ny, nlevs, 1, nm=nm, depth=0.8)
    batch.set_tile(2, hos.HOS_TILE_GROUPS)
    batch.use_graph(graph)
    for i, u in enumerate(states):
        batch.set_state(u, i)
    batch.step(4, dts)
    got = [(batch.get_state(i), batch.energy(i)) for i in range(nm)]
    batch.close()
    for i in (0, 2, nm - 1):
        alone = solver(hos, Nx, Ny, nlevs, 1, depth=0.8)
        alone.set_state(states[i])
        alone.step(4, dts[i])
        assert np.array_equal(alone.get_state(), got[i][0])
        assert alone.energy(0) == got[i][1]
        alone.close()


def test_diag_large_batch_matches_mode_off(hos):
    """Eight 512 x 1024 members, M = 5, FP32 with one row per wave in the x
    passes and two member groups (the large-batch paths): three RK4 steps with
    the diagonal passes stay within the FP32 form's error of the mode off."""
    Nx, Ny, nlevs, nm = 512, 1024, 4, 8
    print("diag large batch", flush=True)
    Lx, Ly = 32 * np.pi, 64 * np.pi
    states = [random_state(Nx, Ny, nlevs, amp=0.03, slope=3.0, seed=90 + i, Lx=Lx, Ly=Ly) for i in range(nm)]
    out = {}
    for diag in (1, 2):
        s = solver(hos, Nx, Ny, nlevs, 1, nm=nm, diag=diag, Lx=Lx, Ly=Ly, depth=1.0)
        s.set_tile(2, hos.HOS_TILE_XPASS)
        s.set_tile(2, hos.HOS_TILE_GROUPS)
        s.use_graph(True)
        for i, u in enumerate(states):
            s.set_state(u, i)
        s.step(3, 0.05)
        out[diag] = [s.get_state(i) for i in (0, nm - 1)]
        s.close()
    for a, b in zip(out[1], out[2]):
        err = rel_l2(b, a)
        print(f"  rel L2 {err:.2e}", flush=True)
        assert err < 1e-6, err


@pytest.mark.parametrize("Nx,Ny", [(512, 64), (1024, 32), (512, 512)])
@pytest.mark.parametrize("nlevs", [1, 2, 4, 7])
@pytest.mark.parametrize("band", [1, 2])
def test_xblock_same_bits(hos, Nx, Ny, nlevs, band):
    """Blocked x-intermediates (HOS_TILE_XBLOCK) change only where the values
    lie: with the diagonal passes and one row per wave in the x passes, RK4
    steps of three members over two member groups, and one RHS, are the same
    bits with the blocks on and off.  Ny = 32 and 64: Nk = 17 and 33 rows, so
    the last block holds one row (and the band of a masked state ends inside
    a block)."""
    print(f"xblock {Nx}x{Ny} nlevs={nlevs} band={band}", flush=True)
    Lx, Ly = 2 * np.pi * Nx / 64, 2 * np.pi * Ny / 64
    nm = 3
    states = [random_state(Nx, Ny, nlevs, amp=0.05, slope=3.0, seed=110 + i, Lx=Lx, Ly=Ly) for i in range(nm)]
    out = {}
    for xb in (1, 2):
        s = solver(hos, Nx, Ny, nlevs, 1, nm=nm, Lx=Lx, Ly=Ly, depth=1.1)
        s.set_tile(2, hos.HOS_TILE_XPASS)
        s.set_tile(2, hos.HOS_TILE_GROUPS)
        s.set_tile(band, hos.HOS_TILE_BAND)
        s.set_tile(xb, hos.HOS_TILE_XBLOCK)
        assert s.get_tile(hos.HOS_TILE_XBLOCK) == xb
        r = s.rhs(states[0])
        for i, u in enumerate(states):
            s.set_state(u, i)
        s.step(3, 0.05)
        out[xb] = (r, [s.get_state(i) for i in range(nm)], [s.energy(i) for i in range(nm)])
        s.close()
    assert np.array_equal(out[1][0], out[2][0])
    for a, b in zip(out[1][1], out[2][1]):
        assert np.array_equal(a, b)
    assert out[1][2] == out[2][2]


@pytest.mark.parametrize("Nx,Ny", [(512, 64), (1024, 128)])
def test_xblock_rhs_vs_oracle(hos, oracle_mod, Nx, Ny):
    """The FP32 diagonal form with wave x passes and blocked x-intermediates
    against the FP64 oracle, finite depth, ramp on."""
    nlevs, depth = 4, 0.9
    Lx, Ly = 2 * np.pi * Nx / 64, 2 * np.pi * Ny / 64
    print(f"xblock rhs {Nx}x{Ny}", flush=True)
    o = oracle_mod.Oracle(Nx, Ny, nlevs, Lx, Ly, depth=depth, rampflg=1, Tramp=2.0)
    s = solver(hos, Nx, Ny, nlevs, 1, Lx=Lx, Ly=Ly, depth=depth, rampflg=1, Tramp=2.0)
    s.set_tile(2, hos.HOS_TILE_XPASS)
    s.set_tile(2, hos.HOS_TILE_XBLOCK)
    u = random_state(Nx, Ny, nlevs, amp=0.05, slope=3.0, seed=31, Lx=Lx, Ly=Ly)
    err = rel_l2(s.rhs(u, 1.3), o.rhs_hos(u, 1.3))
    print(f"  rel L2 {err:.2e}", flush=True)
    assert err < TOL32, err
    s.close()


def test_xblock_bench_shape_same_bits(hos):
    """The benchmark's shape: eight 1024^2 members, M = 5, FP32, kp h = 1, two
    member groups under captured graphs, masked states: three RK4 steps with
    the blocked x-intermediates are the same bits as row by row."""
    N, nlevs, nm = 1024, 4, 8
    L = 64 * np.pi
    print("xblock bench shape", flush=True)
    states = [random_state(N, N, nlevs, amp=0.02, slope=3.0, seed=130 + i, Lx=L, Ly=L) for i in range(nm)]
    out = {}
    for xb in (1, 2):
        s = solver(hos, N, N, nlevs, 1, nm=nm, Lx=L, Ly=L, depth=1.0)
        s.set_tile(2, hos.HOS_TILE_XPASS)
        s.set_tile(2, hos.HOS_TILE_GROUPS)
        s.set_tile(xb, hos.HOS_TILE_XBLOCK)
        s.use_graph(True)
        for i, u in enumerate(states):
            s.set_state(u, i)
        s.step(3, 0.05)
        out[xb] = [s.get_state(i) for i in range(nm)]
        s.close()
    for a, b in zip(out[1], out[2]):
        assert np.array_equal(a, b)
