"""GPU parity: libhos_gpu.so (C ABI) against the FP64 oracle on the same inputs.

Tolerances: FP64 path <= 1e-12 relative L2 for one RHS / transform (both are
FP64 pseudo-spectral evaluations of the same expressions; they differ only in
rounding order), FP32 path <= 2e-6 relative L2 for one RHS (measured 1e-7 to
3e-7 on the random states below; 3e-9 to 6e-9 on the smoother 2048-row ones).
"""
import numpy as np
import pytest

from helpers import dealias_mask, harmonic_state, random_state, to_phys, to_spec, zakharov_rhs
from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 2e-6


def make(hos, Nx, Ny, nlevs, prec, **kw):
    p = hos.params(Nx=Nx, Ny=Ny, nlevs=nlevs, Lx=kw.pop("Lx", 2 * np.pi), Ly=kw.pop("Ly", 2 * np.pi),
                   precision=prec, **kw)
    return hos.Solver(p)


@pytest.mark.parametrize("Nx,Ny", [(16, 16), (64, 32), (32, 128), (256, 256), (2048, 512), (512, 2048)])
@pytest.mark.parametrize("prec", [0, 1])
def test_fft_2d_vs_numpy(hos, Nx, Ny, prec):
    print(f"fft {Nx}x{Ny} prec={prec}", flush=True)
    s = make(hos, Nx, Ny, 2, prec)
    rng = np.random.default_rng(Nx + Ny)
    u = rng.standard_normal((Nx, Ny))
    hu = s.fft_2d(u)
    ref = to_spec(u)
    tol = 1e-14 if prec == 0 else 3e-7
    back = s.ifft_2d(ref)
    print(f"  r2c {rel_l2(hu, ref):.2e} c2r {rel_l2(back, u):.2e}", flush=True)
    assert rel_l2(hu, ref) < tol
    assert rel_l2(back, u) < tol


@pytest.mark.parametrize("nlevs", [0, 1, 2, 3, 4, 5, 7])
@pytest.mark.parametrize("depth", [0.0, 0.7])
@pytest.mark.parametrize("diag", [1, 2])
def test_rhs_fp64_vs_oracle(hos, oracle_mod, nlevs, depth, diag):
    """Both forms of the W expansion: level passes (HOS_TILE_DIAG 1) and
    diagonal passes (2, the default)."""
    Nx, Ny = 64, 32
    print(f"rhs fp64 nlevs={nlevs} depth={depth} diag={diag}", flush=True)
    Lx, Ly = 4 * np.pi, 2 * np.pi
    o = oracle_mod.Oracle(Nx, Ny, nlevs, Lx, Ly, depth=depth)
    s = make(hos, Nx, Ny, nlevs, 0, Lx=Lx, Ly=Ly, depth=depth)
    s.set_tile(diag, hos.HOS_TILE_DIAG)
    u = random_state(Nx, Ny, nlevs, amp=0.08, seed=nlevs)
    ref = o.rhs_hos(u)
    got = s.rhs(u)
    print(f"  rel L2 {rel_l2(got, ref):.2e}", flush=True)
    assert rel_l2(got, ref) < TOL64, rel_l2(got, ref)


@pytest.mark.parametrize("quirk", [0, 1])
@pytest.mark.parametrize("ramp", [0, 1])
def test_rhs_quirk_ramp(hos, oracle_mod, quirk, ramp):
    Nx, Ny, nlevs = 32, 32, 2
    print(f"rhs quirk={quirk} ramp={ramp}", flush=True)
    o = oracle_mod.Oracle(Nx, Ny, nlevs, 2 * np.pi, 2 * np.pi, rampflg=ramp, Tramp=3.0, kx_quirk=bool(quirk))
    s = make(hos, Nx, Ny, nlevs, 0, rampflg=ramp, Tramp=3.0, kx_quirk=quirk)
    u = random_state(Nx, Ny, nlevs, amp=0.1, seed=3)
    for t in (0.0, 1.7):
        assert rel_l2(s.rhs(u, t), o.rhs_hos(u, t)) < TOL64


@pytest.mark.parametrize("Nx,Ny", [(2048, 32), (32, 2048)])
@pytest.mark.parametrize("prec", [0, 1])
def test_rhs_2048_rows_vs_oracle(hos, oracle_mod, Nx, Ny, prec):
    """Rows of 2048 points in x and in y, finite depth: the largest FFT passes
    (the FP64 ones hold more than 64 KiB of LDS per workgroup)."""
    nlevs, depth = 2, 0.9
    Lx, Ly = 2 * np.pi * Nx / 32, 2 * np.pi * Ny / 32
    print(f"rhs {Nx}x{Ny} prec={prec}", flush=True)
    o = oracle_mod.Oracle(Nx, Ny, nlevs, Lx, Ly, depth=depth)
    s = make(hos, Nx, Ny, nlevs, prec, Lx=Lx, Ly=Ly, depth=depth)
    u = random_state(Nx, Ny, nlevs, amp=0.05, slope=3.0, seed=11)
    err = rel_l2(s.rhs(u), o.rhs_hos(u))
    print(f"  rel L2 {err:.2e}", flush=True)
    assert err < (TOL64 if prec == 0 else TOL32), err


@pytest.mark.parametrize("Nx", [512, 1024])
def test_xpass_modes_agree(hos, oracle_mod, Nx):
    """The x inverse passes with rows per workgroup and with one row per wave
    (FP32 form) give the same RHS, both within the FP32 tolerance."""
    Ny, nlevs, depth = 32, 3, 0.8
    Lx, Ly = 2 * np.pi * Nx / 32, 2 * np.pi
    print(f"x-pass modes {Nx}x{Ny}", flush=True)
    o = oracle_mod.Oracle(Nx, Ny, nlevs, Lx, Ly, depth=depth)
    s = make(hos, Nx, Ny, nlevs, 1, Lx=Lx, Ly=Ly, depth=depth)
    u = random_state(Nx, Ny, nlevs, amp=0.05, slope=3.0, seed=5)
    ref = o.rhs_hos(u)
    s.set_tile(1, hos.HOS_TILE_XPASS)
    assert s.get_tile(hos.HOS_TILE_XPASS) == 1
    a = s.rhs(u)
    s.set_tile(2, hos.HOS_TILE_XPASS)
    assert s.get_tile(hos.HOS_TILE_XPASS) == 2
    b = s.rhs(u)
    print(f"  rows per workgroup {rel_l2(a, ref):.2e}, row per wave {rel_l2(b, ref):.2e}", flush=True)
    assert rel_l2(a, ref) < TOL32 and rel_l2(b, ref) < TOL32, (rel_l2(a, ref), rel_l2(b, ref))
    assert rel_l2(a, b) < 1e-6


@pytest.mark.parametrize("nlevs", [1, 2, 4, 5, 7])
@pytest.mark.parametrize("diag", [1, 2])
def test_rhs_fp32_vs_oracle(hos, oracle_mod, nlevs, diag):
    """FP32 RHS for M = 2..8, both W-expansion forms: with the level passes,
    up to M = 5 the RHS pass reads its physical fields before the deferred
    ones, above that after them."""
    Nx, Ny = 128, 64
    print(f"rhs fp32 nlevs={nlevs} diag={diag}", flush=True)
    o = oracle_mod.Oracle(Nx, Ny, nlevs, 2 * np.pi, np.pi)
    s = make(hos, Nx, Ny, nlevs, 1, Ly=np.pi)
    s.set_tile(diag, hos.HOS_TILE_DIAG)
    u = random_state(Nx, Ny, nlevs, amp=0.05, seed=7)
    err = rel_l2(s.rhs(u), o.rhs_hos(u))
    print(f"  rel L2 {err:.2e}", flush=True)
    assert err < TOL32


@pytest.mark.parametrize("nlevs", [1, 2, 3, 4])
def test_zvel_vs_oracle(hos, oracle_mod, nlevs):
    Nx, Ny = 32, 64
    print(f"zvel nlevs={nlevs}", flush=True)
    o = oracle_mod.Oracle(Nx, Ny, nlevs, 2 * np.pi, 2 * np.pi)
    s = make(hos, Nx, Ny, nlevs, 0)
    u = random_state(Nx, Ny, nlevs, amp=0.1, seed=11)
    ref = o.Zvel(u)
    got = s.zvel(u)
    for r, g in zip(ref, got):
        if np.linalg.norm(r) == 0:
            assert np.linalg.norm(g) < 1e-14
        else:
            assert rel_l2(g, r) < TOL64


@pytest.mark.parametrize("depth", [0.0, 0.8])
def test_truncation_order_in_amplitude(hos, depth):
    """The FP64 GPU path on its own (no oracle) against exact harmonic
    potentials scaled by eps: W^(M), (W^2)^(M) and both RHS components miss the
    exact values (the Zakharov form with exact W for the RHS) by
    O(eps^(M+1)), W^(M-2) by O(eps^max(M-1,1)) and (W^2)^(M-2) by
    O(eps^max(M-1,2)).  The error slopes between eps = 1/4, 1/8 and 1/16 must
    match those orders: a term of the fused algorithm dropped or mis-weighted
    at order n leaves a slope of n."""
    Nx = Ny = 64
    L = 2 * np.pi
    g = 9.81
    for nlevs in range(0, 5):
        M = nlevs + 1
        print(f"truncation order depth={depth} M={M}", flush=True)
        s = make(hos, Nx, Ny, nlevs, 0, depth=depth, g=g)
        keep = ~dealias_mask(Nx, Ny, nlevs)
        errs = []
        for eps in (0.25, 0.125, 0.0625):
            u, W, eta, phi = harmonic_state(Nx, Ny, L, depth, amp_eta=0.08 * eps, amp_phi=0.3 * eps)
            wM, wM2, w2M, w2M2 = (to_phys(w, Nx, Ny) for w in s.zvel(u))
            et, pt = zakharov_rhs(eta, phi, W, L, g)
            r = s.rhs(u)
            errs.append([np.abs(wM - W).max(), np.abs(wM2 - W).max(), np.abs(w2M - W * W).max(),
                         np.abs(w2M2 - W * W).max(), np.abs(r[0] - to_spec(et) * keep).max(),
                         np.abs(r[1] - to_spec(pt) * keep).max()])
        s.close()
        errs = np.array(errs)
        slopes = np.log2(errs[:-1] / errs[1:])
        want = np.array([M + 1, max(M - 1, 1), M + 1, max(M - 1, 2), M + 1, M + 1])
        assert np.all(np.abs(slopes - want) < 0.15), (M, slopes, want)


@pytest.mark.parametrize("scheme", [1, 2, 3, 4, 5])
def test_steps_vs_oracle(hos, oracle_mod, scheme):
    Nx, Ny, nlevs = 32, 32, 2
    print(f"steps scheme={scheme}", flush=True)
    o = oracle_mod.Oracle(Nx, Ny, nlevs, 2 * np.pi, 2 * np.pi, scheme=scheme)
    s = make(hos, Nx, Ny, nlevs, 0, scheme=scheme)
    u = random_state(Nx, Ny, nlevs, amp=0.05, seed=5)
    s.set_state(u)
    dt = 0.05
    ur = u.copy()
    t = 0.0
    Hs = []
    for n in range(10):
        t, H = o.step(ur, t, dt)
        Hs.append(H)
    s.step(10, dt)
    got = s.get_state()
    assert rel_l2(got, ur) < 1e-12
    Hg, vol = s.energy()
    assert abs(Hg - Hs[-1]) <= 1e-12 * abs(Hs[-1])
    assert abs(s.get_time() - t) < 1e-12


def test_api_misuse_is_an_error_and_leaves_the_state(hos):
    """Bad member indices, active counts, tiles and step counts return
    HOS_E_INVALID (HosError) and leave the members' state and time as they were;
    while a step_async is in flight, calls that write the state, the time, the
    tiles or the shared scratch return HOS_E_STATE and change nothing."""
    Nx, Ny, nlevs, nm = 32, 32, 2, 2
    print("api misuse", flush=True)
    p = hos.params(Nx=Nx, Ny=Ny, nlevs=nlevs, precision=1)
    s = hos.Solver(p, nm)
    u = random_state(Nx, Ny, nlevs, amp=0.05, seed=9)
    s.set_state(u, 0)
    s.set_state(u, 1)
    s.step(2, 0.02)
    before = [s.get_state(i) for i in range(nm)], [s.get_time(i) for i in range(nm)]
    bad = [
        lambda: s.set_state(u, nm),
        lambda: s.set_state(u, -1),
        lambda: s.get_state(nm),
        lambda: s.set_active(nm + 1),
        lambda: s.set_active(-1),
        lambda: s.copy_member(0, nm),
        lambda: s.set_member(nm, 0.0, 0.02),
        lambda: s.set_tile(5, hos.HOS_TILE_C2R),
        lambda: s.set_tile(3, hos.HOS_TILE_XPASS),
        lambda: s.set_tile(5, hos.HOS_TILE_GROUPS),  # 1 .. 4 groups
        lambda: s.step(-1, 0.02),
    ]
    for f in bad:
        with pytest.raises(hos.HosError):
            f()
    after = [s.get_state(i) for i in range(nm)], [s.get_time(i) for i in range(nm)]
    for a, b in zip(before[0], after[0]):
        assert np.array_equal(a, b)
    assert before[1] == after[1]
    s.step_async(1)
    for f in (lambda: s.step(1, 0.02), lambda: s.set_state(u, 0), lambda: s.rhs(u),
              lambda: s.set_tile(8, hos.HOS_TILE_C2R), lambda: s.set_active(1)):
        with pytest.raises(hos.HosError):
            f()
    s.wait()
    t1 = s.get_time(0)
    assert t1 == pytest.approx(before[1][0] + 0.02)  # the step in flight kept its dt
    s.step(1, 0.02)  # usable again
    s.close()


def test_members_match_alone(hos):
    """Each member of a batch gives the result it gives alone (bit-identical)."""
    Nx, Ny, nlevs = 64, 64, 3
    print("members", flush=True)
    p = hos.params(Nx=Nx, Ny=Ny, nlevs=nlevs, precision=1)
    states = [random_state(Nx, Ny, nlevs, amp=0.06, seed=20 + i) for i in range(3)]
    batch = hos.Solver(p, 3)
    for i, u in enumerate(states):
        batch.set_state(u, i)
    dts = np.array([0.02, 0.03, 0.025])
    batch.step(5, dts)
    for i, u in enumerate(states):
        alone = hos.Solver(p, 1)
        alone.set_state(u)
        alone.step(5, dts[i])
        assert np.array_equal(alone.get_state(), batch.get_state(i))


def test_many_rows_batch_matches_alone(hos):
    """With at least 4096 spectral rows in flight (8 members at 512 x 1024)
    the batch takes its large-batch paths: the x inverse outputs in one group
    per source, non-temporal stores in the wave x passes and non-temporal loads
    in the RK epilogue, over two member groups.  Members stepped that way are
    bit-identical to the same members stepped alone (one member: 513 rows,
    outputs in groups of 3, plain loads), both with one row per wave."""
    Nx, Ny, nlevs, nm = 512, 1024, 4, 8
    print("many rows batch", flush=True)
    p = hos.params(Nx=Nx, Ny=Ny, nlevs=nlevs, precision=1, depth=1.0, Lx=32 * np.pi, Ly=64 * np.pi)
    states = [random_state(Nx, Ny, nlevs, amp=0.03, slope=3.0, seed=60 + i, Lx=32 * np.pi, Ly=64 * np.pi)
              for i in range(nm)]
    dts = 0.02 + 0.001 * np.arange(nm)
    batch = hos.Solver(p, nm)
    batch.set_tile(2, hos.HOS_TILE_XPASS)
    batch.set_tile(2, hos.HOS_TILE_GROUPS)
    batch.use_graph(True)
    for i, u in enumerate(states):
        batch.set_state(u, i)
    batch.step(3, dts)
    got = [(batch.get_state(i), batch.energy(i)) for i in (0, nm - 1)]
    batch.close()
    for (u_b, e_b), i in zip(got, (0, nm - 1)):
        alone = hos.Solver(p, 1)
        alone.set_tile(2, hos.HOS_TILE_XPASS)
        alone.set_state(states[i])
        alone.step(3, dts[i])
        assert np.array_equal(alone.get_state(), u_b)
        assert alone.energy(0) == e_b
        alone.close()


@pytest.mark.parametrize("graph", [False, True])
def test_member_groups_match_one_chain(hos, graph):
    """Member groups on their own streams (2, 3, 4) give bit-identical states
    and energies to one launch chain, for uneven ranges, and after the active
    count drops (fewer groups) and comes back."""
    Nx, Ny, nlevs, nm = 64, 32, 3, 9
    print(f"member groups graph={graph}", flush=True)
    p = hos.params(Nx=Nx, Ny=Ny, nlevs=nlevs, precision=1, depth=0.9)
    states = [random_state(Nx, Ny, nlevs, amp=0.06, seed=40 + i) for i in range(nm)]
    dts = 0.02 + 0.001 * np.arange(nm)
    out = {}
    for ng in (1, 2, 3, 4):
        s = hos.Solver(p, nm)
        s.set_tile(ng, hos.HOS_TILE_GROUPS)
        assert s.get_tile(hos.HOS_TILE_GROUPS) == ng
        s.use_graph(graph)
        for i, u in enumerate(states):
            s.set_state(u, i)
        s.step(3, dts)
        s.set_active(5)  # at most 2 groups of two members or more
        s.step(2, dts)
        s.set_active(nm)
        s.step(3, dts)
        out[ng] = [(s.get_state(i), s.energy(i)) for i in range(nm)]
        s.close()
    for ng in (2, 3, 4):
        for (u1, e1), (u2, e2) in zip(out[1], out[ng]):
            assert np.array_equal(u1, u2)
            assert e1 == e2


def test_slots_active_count_and_member_params(hos):
    """Per-member depth/dt, stepping a prefix of the slots, slot copies, and
    two contexts stepped concurrently (async): each member matches its run
    alone bit for bit."""
    Nx, Ny, nlevs = 64, 32, 2
    print("slots", flush=True)
    p = hos.params(Nx=Nx, Ny=Ny, nlevs=nlevs, precision=1, Lx=8.0, Ly=4.0)
    depths, dts = [0.0, 0.6, 1.4, 0.0], [0.02, 0.025, 0.03, 0.015]
    states = [random_state(Nx, Ny, nlevs, amp=0.05, seed=40 + i) for i in range(4)]
    pool = hos.Solver(p, 4)
    other = hos.Solver(p, 1)
    for i in range(4):
        pool.set_state(states[i], i)
        pool.set_member(i, depths[i], dts[i])
    other.set_state(states[3])
    other.set_member(0, depths[3], dts[3])
    pool.use_graph(True)
    pool.set_active(2)
    pool.step_async(3)
    other.step_async(3)
    pool.wait()
    other.wait()
    assert np.array_equal(pool.get_state(2), states[2])  # not active: untouched
    for i in range(2):
        alone = hos.Solver(hos.params(Nx=Nx, Ny=Ny, nlevs=nlevs, precision=1, Lx=8.0, Ly=4.0, depth=depths[i]))
        alone.set_state(states[i])
        alone.step(3, dts[i])
        assert np.array_equal(alone.get_state(), pool.get_state(i))
        assert alone.get_time() == pytest.approx(pool.get_time(i))
    pool.copy_member(1, 3)
    ref1 = pool.get_state(1)
    assert np.array_equal(pool.get_state(3), ref1)
    pool.set_active(4)
    pool.step_async(2)
    pool.wait()
    assert np.array_equal(pool.get_state(3), pool.get_state(1))


def test_autotune_restores_state_and_tiles_agree(hos):
    """The autotuner leaves state and time as they were, and every y tile
    gives the same result (tiles change launch shape, not arithmetic)."""
    Nx, Ny, nlevs = 64, 64, 3
    print("autotune", flush=True)
    p = hos.params(Nx=Nx, Ny=Ny, nlevs=nlevs, precision=1)
    u = random_state(Nx, Ny, nlevs, amp=0.06, seed=31)
    s = hos.Solver(p, 2)
    s.set_state(u, 0)
    s.set_state(u, 1)
    s.set_member(0, 0.0, 0.02)
    s.set_member(1, 0.0, 0.035)
    s.step_async(1)  # per-member dt
    s.wait()
    u1 = s.get_state(0)
    e_before = [s.energy(i) for i in range(2)]
    t_before = [s.get_time(i) for i in range(2)]
    tiles = s.autotune(2)
    assert all(t in (4, 8, 16) for t in tiles)
    assert np.array_equal(s.get_state(0), u1)
    assert [s.get_time(i) for i in range(2)] == t_before
    assert [s.energy(i) for i in range(2)] == e_before  # H and volume restored
    s.step_async(1)  # the per-member dt restored on the device
    s.wait()
    assert s.get_time(0) == pytest.approx(t_before[0] + 0.02, abs=1e-15)
    assert s.get_time(1) == pytest.approx(t_before[1] + 0.035, abs=1e-15)
    outs = []
    for tx in (4, 8, 16):
        r = hos.Solver(p)
        r.set_state(u)
        r.set_tile(tx)
        r.step(3, 0.02)
        outs.append(r.get_state())
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


@pytest.mark.parametrize("scheme", [1, 2, 3, 4, 5])
def test_mean_level_correction_and_volume(hos, scheme):
    """mean_level=1 holds eta-hat(0,0) at its initial value (FP64) with every
    time scheme (the generic Butcher path of Dormand-Prince included); the
    reported volume is Re eta-hat(0,0) at the start of the last step."""
    Nx, Ny, nlevs = 64, 64, 3
    print(f"mean level scheme={scheme}", flush=True)
    u = random_state(Nx, Ny, nlevs, amp=0.08, seed=17)
    u[0, 0, 0] = 0.0123
    for ml in (0, 1):
        s = hos.Solver(hos.params(Nx=Nx, Ny=Ny, nlevs=nlevs, precision=1, mean_level=ml, scheme=scheme))
        s.set_state(u)
        s.step(5, 0.02)
        st = s.get_state()
        if ml:
            assert st[0, 0, 0] == 0.0123
        H, vol = s.energy()
        assert abs(vol - 0.0123) < 1e-6
        s.step(1, 0.02)
        assert s.energy()[1] == st[0, 0, 0].real


def test_graph_replay_matches_eager(hos):
    Nx, Ny, nlevs = 64, 64, 4
    print("graph", flush=True)
    p = hos.params(Nx=Nx, Ny=Ny, nlevs=nlevs, precision=1)
    u = random_state(Nx, Ny, nlevs, amp=0.06, seed=9)
    a = hos.Solver(p)
    b = hos.Solver(p)
    a.set_state(u)
    b.set_state(u)
    b.use_graph(True)
    for rep in range(3):
        a.step(4, 0.02)
        b.step(4, 0.02)
        assert np.array_equal(a.get_state(), b.get_state())


@pytest.mark.parametrize("prec,Nx,Ny,nlevs,nm,xmode", [
    (1, 1024, 512, 4, 4, 2),   # wave x passes, two member groups
    (1, 512, 256, 3, 2, 1),    # two rows per x workgroup
    (0, 256, 128, 2, 2, 1),    # FP64, four rows per x workgroup
    (1, 128, 256, 5, 3, 1),    # eight rows per x workgroup
])
def test_band_of_masked_states_is_bit_identical(hos, prec, Nx, Ny, nlevs, nm, xmode):
    """Masked states (the mask of operators_par.c:309-322 leaves their other
    modes zero) step with the band on (the level-0 pass, the y passes that
    read it, the RHS pass output and the RK epilogue skip the masked modes)
    bit-identically to the band off, states and energies alike, with finite
    depth and the ramp on; a member whose state is not masked turns the band
    off for its batch, and it then matches a context with the band off."""
    print(f"band prec={prec} {Nx}x{Ny} nlevs={nlevs} members={nm}", flush=True)
    Lx, Ly = 2 * np.pi * Nx / 32, 2 * np.pi * Ny / 32
    p = hos.params(Nx=Nx, Ny=Ny, nlevs=nlevs, Lx=Lx, Ly=Ly, precision=prec, depth=1.1, rampflg=1, Tramp=0.5)
    states = [random_state(Nx, Ny, nlevs, amp=0.04, slope=3.0, seed=70 + i, Lx=Lx, Ly=Ly) for i in range(nm)]
    dts = 0.02 + 0.002 * np.arange(nm)
    out = {}
    for band in (1, 2):
        s = hos.Solver(p, nm)
        s.set_tile(band, hos.HOS_TILE_BAND)
        s.set_tile(xmode, hos.HOS_TILE_XPASS)
        s.use_graph(band == 2)
        for i, u in enumerate(states):
            s.set_state(u, i)
        s.step(4, dts)
        out[band] = [(s.get_state(i), s.energy(i)) for i in range(nm)]
        s.close()
    for (u1, e1), (u2, e2) in zip(out[1], out[2]):
        assert np.array_equal(u1, u2)
        assert e1 == e2
    # the masked modes stay exactly zero
    cut = dealias_mask(Nx, Ny, nlevs)
    assert not np.any(out[2][0][0][:, cut])
    # an unmasked member: the batch runs without the band, as the band-off context
    bad = states[0].copy()
    bad[0][cut] = 1e-9
    res = []
    for band in (1, 2):
        s = hos.Solver(p, nm)
        s.set_tile(band, hos.HOS_TILE_BAND)
        s.set_tile(xmode, hos.HOS_TILE_XPASS)
        for i, u in enumerate(states):
            s.set_state(bad if i == 0 else u, i)
        s.step(2, dts)
        res.append([s.get_state(i) for i in range(nm)])
        s.close()
    for a, b in zip(*res):
        assert np.array_equal(a, b)
    assert np.array_equal(res[1][0][0][cut], bad[0][cut])  # its masked modes kept their values


@pytest.mark.parametrize("prec,Nx,Ny,nlevs,nm,c2r", [
    (1, 1024, 512, 5, 4, 8),   # wave y passes, three fused levels, two member groups
    (1, 256, 1024, 6, 2, 4),   # four fused levels, four x-rows per y tile
    (1, 128, 256, 4, 3, 16),   # sixteen x-rows per y tile
    (0, 256, 128, 5, 2, 8),    # FP64: the workgroup FFT in the y c2r
])
def test_fused_level_passes_are_bit_identical(hos, prec, Nx, Ny, nlevs, nm, c2r):
    """A W level's y c2r launch that also does the next level's y r2c
    (HOS_TILE_FUSE on) steps states and energies bit-identically to the
    separate launches (off) with the same tiles, RHS and Zvel alike (the level
    passes: HOS_TILE_DIAG off)."""
    print(f"fuse prec={prec} {Nx}x{Ny} nlevs={nlevs} members={nm} c2r={c2r}", flush=True)
    Lx, Ly = 2 * np.pi * Nx / 32, 2 * np.pi * Ny / 32
    p = hos.params(Nx=Nx, Ny=Ny, nlevs=nlevs, Lx=Lx, Ly=Ly, precision=prec, depth=1.3)
    states = [random_state(Nx, Ny, nlevs, amp=0.05, slope=3.0, seed=90 + i, Lx=Lx, Ly=Ly) for i in range(nm)]
    out = {}
    for fuse in (1, 2):
        s = hos.Solver(p, nm)
        s.set_tile(1, hos.HOS_TILE_DIAG)
        s.set_tile(fuse, hos.HOS_TILE_FUSE)
        assert s.get_tile(hos.HOS_TILE_FUSE) == fuse
        s.set_tile(c2r, hos.HOS_TILE_ALL)
        s.use_graph(fuse == 2)
        for i, u in enumerate(states):
            s.set_state(u, i)
        r = s.rhs(states[0])
        z = s.zvel(states[1 % nm])
        s.step(3, 0.02)
        out[fuse] = (r, z, [(s.get_state(i), s.energy(i)) for i in range(nm)])
        s.close()
    assert np.array_equal(out[1][0], out[2][0])
    for a, b in zip(out[1][1], out[2][1]):
        assert np.array_equal(a, b)
    for (u1, e1), (u2, e2) in zip(out[1][2], out[2][2]):
        assert np.array_equal(u1, u2)
        assert e1 == e2


@pytest.mark.parametrize("prec,ng,graph", [(1, 2, True), (1, 3, False), (0, 2, True)])
def test_offset_member_groups_are_bit_identical(hos, prec, ng, graph):
    """Member groups that run free over a call, the others offset behind
    group 0 (HOS_TILE_SKEW on), step every member, states, energies and
    times, as the groups joined after every step (off); the steps are timed
    one by one either way, and calls of one step and of several agree."""
    Nx, Ny, nlevs, nm = 128, 256, 4, 2 * ng + 1
    print(f"skew prec={prec} groups={ng} graph={graph}", flush=True)
    Lx, Ly = 2 * np.pi * 4, 2 * np.pi * 8
    p = hos.params(Nx=Nx, Ny=Ny, nlevs=nlevs, Lx=Lx, Ly=Ly, precision=prec, depth=1.4)
    states = [random_state(Nx, Ny, nlevs, amp=0.05, slope=3.0, seed=110 + i, Lx=Lx, Ly=Ly) for i in range(nm)]
    dts = 0.02 + 0.001 * np.arange(nm)
    out = {}
    for skew in (1, 2):
        s = hos.Solver(p, nm)
        s.set_tile(ng, hos.HOS_TILE_GROUPS)
        s.set_tile(skew, hos.HOS_TILE_SKEW)
        assert s.get_tile(hos.HOS_TILE_SKEW) == skew
        s.use_graph(graph)
        for i, u in enumerate(states):
            s.set_state(u, i)
        s.step(5, dts)
        st = s.step_times()
        assert len(st) == 5 and all(t > 0 for t in st)
        s.step(1, dts)
        s.step(3, dts)
        out[skew] = [(s.get_state(i), s.energy(i), s.get_time(i)) for i in range(nm)]
        s.close()
    for (u1, e1, t1), (u2, e2, t2) in zip(out[1], out[2]):
        assert np.array_equal(u1, u2)
        assert e1 == e2
        assert t1 == t2


@pytest.mark.parametrize("prec", [0, 1])
def test_rhs_fft_plans_vs_oracle(hos, oracle_mod, prec):
    """Every FFT radix plan of the RHS pass (HOS_TILE_PLAN 1-4: the built-in
    rule, radices of at most 8, radix 4, the largest radix at every stage)
    gives the oracle's RHS within the tolerance of its precision; the plans
    other than the built-in one are refused below Ny = 512."""
    Nx, Ny, nlevs, depth = 64, 512, 4, 0.9
    Lx, Ly = 2 * np.pi * 2, 2 * np.pi * 16
    print(f"rhs fft plans prec={prec}", flush=True)
    o = oracle_mod.Oracle(Nx, Ny, nlevs, Lx, Ly, depth=depth)
    s = make(hos, Nx, Ny, nlevs, prec, Lx=Lx, Ly=Ly, depth=depth)
    u = random_state(Nx, Ny, nlevs, amp=0.05, slope=3.0, seed=19, Lx=Lx, Ly=Ly)
    ref = o.rhs_hos(u)
    for plan in (1, 2, 3, 4):
        s.set_tile(plan, hos.HOS_TILE_PLAN)
        assert s.get_tile(hos.HOS_TILE_PLAN) == plan
        err = rel_l2(s.rhs(u), ref)
        print(f"  plan {plan}: rel L2 {err:.2e}", flush=True)
        assert err < (TOL64 if prec == 0 else TOL32), (plan, err)
    s.close()
    small = make(hos, 64, 256, 2, prec)
    with pytest.raises(hos.HosError):
        small.set_tile(2, hos.HOS_TILE_PLAN)
    small.set_tile(1, hos.HOS_TILE_PLAN)
    small.close()
