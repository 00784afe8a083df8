"""Pins of the FP64 oracle (oracle/hos_oracle.c) against facts that do not
come from the oracle itself: brute-force DFT, numpy's FFT, trigonometric
identities, exact harmonic potentials, Runge-Kutta order conditions,
Parseval, and the Stokes-wave phase speed.  No GPU."""
import numpy as np
import pytest

from conftest import rel_l2
from helpers import dealias_mask, grid, harmonic_state, random_state, to_phys, to_spec, zakharov_rhs


def brute_r2c(u):
    """O(N^2) DFT in long double: hu[ky][kx] = 1/(NxNy) sum u e^{-i(...)}."""
    Nx, Ny = u.shape
    x = np.arange(Nx, dtype=np.longdouble)
    y = np.arange(Ny, dtype=np.longdouble)
    out = np.zeros((Ny // 2 + 1, Nx), complex)
    for ky in range(Ny // 2 + 1):
        for kx in range(Nx):
            ph = -2 * np.pi * (np.outer(x * kx / Nx, np.ones(Ny)) + np.outer(np.ones(Nx), y * ky / Ny))
            re = np.sum(u.astype(np.longdouble) * np.cos(ph))
            im = np.sum(u.astype(np.longdouble) * np.sin(ph))
            out[ky, kx] = complex(float(re), float(im)) / (Nx * Ny)
    return out


# ---------------------------------------------------------------- FFT (A6)
@pytest.mark.parametrize("Nx,Ny", [(8, 6), (16, 16), (12, 10)])
def test_fft_vs_bruteforce_dft(oracle_mod, Nx, Ny):
    o = oracle_mod.Oracle(Nx, Ny, 2, 1.0, 1.0)
    u = np.random.default_rng(1).standard_normal((Nx, Ny))
    assert np.abs(o.fft_2d(u) - brute_r2c(u)).max() < 1e-15


@pytest.mark.parametrize("Nx,Ny", [(64, 32), (32, 128), (256, 64)])
def test_fft_vs_numpy_and_roundtrip(oracle_mod, Nx, Ny):
    o = oracle_mod.Oracle(Nx, Ny, 2, 1.0, 1.0)
    u = np.random.default_rng(2).standard_normal((Nx, Ny))
    hu = o.fft_2d(u)
    assert rel_l2(hu, to_spec(u)) < 1e-15
    assert np.abs(o.ifft_2d(hu) - u).max() < 1e-13
    assert hu[0, 0].real == pytest.approx(u.mean(), abs=1e-15)  # eta-hat(0,0) is the mean


def test_c2r_uses_hermitian_projection(oracle_mod):
    """A ky=0 column that is not Hermitian in kx is inverted as its Hermitian
    projection (the stated policy; FFTW leaves this case undefined)."""
    Nx, Ny = 16, 8
    o = oracle_mod.Oracle(Nx, Ny, 2, 1.0, 1.0)
    hu = np.zeros((Ny // 2 + 1, Nx), complex)
    hu[0, 3] = 1.0 + 0.5j            # no partner at kx = -3
    hu[2, 5] = 0.25 - 0.75j          # interior column: always fine
    proj = hu.copy()
    proj[0, 3] = 0.5 * (1.0 + 0.5j)
    proj[0, Nx - 3] = np.conj(0.5 * (1.0 + 0.5j))
    ref = np.fft.irfft2(proj.T * Nx * Ny, s=(Nx, Ny))
    assert np.abs(o.ifft_2d(hu) - ref).max() < 1e-14


# ---------------------------------------------------------------- Dx Dy Dz (A7)
def test_derivatives_of_single_modes(oracle_mod):
    Nx, Ny, Lx, Ly = 32, 16, 3.0, 2.0
    o = oracle_mod.Oracle(Nx, Ny, 2, Lx, Ly)
    x, y = grid(Nx, Ny, Lx, Ly)
    for (m, n) in [(3, 2), (-5, 1), (7, 0), (-2, 4)]:
        kx, ky = 2 * np.pi * m / Lx, 2 * np.pi * n / Ly
        f = np.cos(kx * x + ky * y + 0.3)
        hf = o.fft_2d(f)
        assert np.abs(o.ifft_2d(o.Dx(hf)) + kx * np.sin(kx * x + ky * y + 0.3)).max() < 1e-12
        assert np.abs(o.ifft_2d(o.Dy(hf)) + ky * np.sin(kx * x + ky * y + 0.3)).max() < 1e-12
        # d/dz of the deep-water harmonic e^{|k| z} cos(...) at z = 0
        assert np.abs(o.ifft_2d(o.Dz(hf)) - np.hypot(kx, ky) * f).max() < 1e-12


def test_reference_kx_table_quirk(oracle_mod):
    """operators_par.c:55: index i = Nx/2+1 gets +(Nx/2+1) Kx0 (not -(Nx/2-1))."""
    Nx, Ny = 16, 8
    o = oracle_mod.Oracle(Nx, Ny, 2, 2 * np.pi, 2 * np.pi, kx_quirk=True)
    s = oracle_mod.Oracle(Nx, Ny, 2, 2 * np.pi, 2 * np.pi, kx_quirk=False)
    hu = np.zeros((Ny // 2 + 1, Nx), complex)
    hu[1, Nx // 2 + 1] = 1.0
    assert o.Dx(hu)[1, Nx // 2 + 1] == pytest.approx(1j * (Nx // 2 + 1))
    assert s.Dx(hu)[1, Nx // 2 + 1] == pytest.approx(-1j * (Nx // 2 - 1))


def test_quirk_invisible_after_mask(oracle_mod):
    """For NLevs >= 1 the reference kx table and the signed one give the same
    dealiased RHS (the affected modes never re-enter the kept band)."""
    for nlevs in (1, 2, 3):
        Nx, Ny = 32, 32
        u = random_state(Nx, Ny, nlevs, amp=0.1, seed=nlevs)
        a = oracle_mod.Oracle(Nx, Ny, nlevs, 2 * np.pi, 2 * np.pi, kx_quirk=True).rhs_hos(u)
        b = oracle_mod.Oracle(Nx, Ny, nlevs, 2 * np.pi, 2 * np.pi, kx_quirk=False).rhs_hos(u)
        assert rel_l2(a, b) < 1e-13


# ---------------------------------------------------------------- Dealias (A8)
@pytest.mark.parametrize("N,nlevs,mx", [(256, 2, 64), (1024, 4, 170), (512, 3, 102), (64, 0, 32)])
def test_dealias_band(oracle_mod, N, nlevs, mx):
    assert oracle_mod.dealias_mx(N, nlevs) == mx


def test_dealias_mask_symmetric_band(oracle_mod):
    Nx, Ny, nlevs = 32, 16, 2
    o = oracle_mod.Oracle(Nx, Ny, nlevs, 1.0, 1.0)
    kept = o.Dealias(np.ones((Ny // 2 + 1, Nx), complex)) != 0
    mx, my = oracle_mod.dealias_mx(Nx, nlevs), oracle_mod.dealias_mx(Ny, nlevs)
    sk = np.fft.fftfreq(Nx, 1.0 / Nx)
    for j in range(Ny // 2 + 1):
        for i in range(Nx):
            assert kept[j, i] == (abs(sk[i]) <= mx - 1 and j <= my - 1)
    assert np.array_equal(~kept, dealias_mask(Nx, Ny, nlevs))


# ---------------------------------------------------------------- Mult (A5)
def test_mult_single_modes_with_aliasing(oracle_mod):
    Nx, Ny, L = 16, 16, 2 * np.pi
    o = oracle_mod.Oracle(Nx, Ny, 2, L, L)
    x, y = grid(Nx, Ny, L, L)
    a = np.cos(3 * x + 2 * y)
    b = np.cos(6 * x - 1 * y)
    prod = o.ifft_2d(o.Mult(o.fft_2d(a), o.fft_2d(b)))
    # cos A cos B = (cos(A+B) + cos(A-B))/2 ; 9x wraps to -7x on a 16-point grid
    ref = 0.5 * (np.cos(-7 * x + y) + np.cos(-3 * x + 3 * y))
    assert np.abs(prod - ref).max() < 5e-14
    # in place (hu1 = hu2 = hprod) is allowed by the reference
    ha = o.fft_2d(a)
    assert np.abs(o.ifft_2d(o.Mult(ha, ha)) - a * a).max() < 1e-14


# ---------------------------------------------------------------- Zvel (A4)
@pytest.mark.parametrize("depth", [0.0, 0.8])
def test_zvel_converges_to_exact_vertical_velocity(oracle_mod, depth):
    """W^(M) -> dPhi/dz(z=eta) with error falling by a factor ~k*eta per order."""
    Nx = Ny = 64
    L = 2 * np.pi
    u, W, eta, _ = harmonic_state(Nx, Ny, L, depth)
    errs = []
    for nlevs in range(0, 6):
        o = oracle_mod.Oracle(Nx, Ny, nlevs, L, L, depth=depth)
        wM, wM2, w2M, w2M2 = o.Zvel(u)
        errs.append(np.abs(o.ifft_2d(wM) - W).max() / np.abs(W).max())
        if nlevs >= 2:
            wm2 = o.ifft_2d(wM2)
            assert np.abs(wm2 - W).max() / np.abs(W).max() < 3 * max(errs[nlevs - 2], 1e-12)
        if nlevs >= 1:
            w2 = o.ifft_2d(w2M)
            assert np.abs(w2 - W * W).max() / np.abs(W * W).max() < 5 * errs[nlevs - 1] + 1e-12
    for m in range(5):
        assert errs[m + 1] < 0.45 * errs[m], errs
    assert errs[5] < 1e-5, errs


def test_zvel_closed_forms(oracle_mod):
    Nx = Ny = 32
    o = oracle_mod.Oracle(Nx, Ny, 3, 2 * np.pi, 2 * np.pi)
    u = random_state(Nx, Ny, 3, amp=0.1, seed=4)
    flat = u.copy()
    flat[0] = 0  # eta = 0: W^(M) = |k| phi-hat exactly, higher orders vanish
    wM, wM2, w2M, w2M2 = o.Zvel(flat)
    assert np.abs(wM - o.Dz(flat[1])).max() < 1e-15
    still = u.copy()
    still[1] = 0  # phi = 0: W = 0
    assert all(np.abs(w).max() == 0 for w in o.Zvel(still))


# ---------------------------------------------------------------- rhs_hos (A3)
@pytest.mark.parametrize("nlevs", [0, 1, 2, 4])
def test_rhs_single_mode_closed_form(oracle_mod, nlevs):
    Nx = Ny = 32
    L = 2 * np.pi
    o = oracle_mod.Oracle(Nx, Ny, nlevs, L, L)
    x, y = grid(Nx, Ny, L, L)
    k, b = 2.0, 0.1  # 2k stays inside the kept band up to NLevs = 4 at N = 32
    u = np.zeros((2, Ny // 2 + 1, Nx), complex)
    u[1] = to_spec(b * np.cos(k * x) + 0 * y)
    r = o.rhs_hos(u)
    assert np.abs(o.ifft_2d(r[0]) - k * b * np.cos(k * x)).max() < 1e-14
    if nlevs >= 1:  # 1/2 W^2 - 1/2 |grad phi|^2
        ref = 0.5 * k * k * b * b * np.cos(2 * k * x)
    else:           # (W^2)^(M) is empty at NLevs = 0
        ref = -0.5 * k * k * b * b * np.sin(k * x) ** 2
        ref = to_phys(o.Dealias(to_spec(ref + 0 * y)), Nx, Ny)
    assert np.abs(o.ifft_2d(r[1]) - ref).max() < 1e-14


def test_rhs_lake_at_rest_and_still_potential(oracle_mod):
    Nx = Ny = 32
    o = oracle_mod.Oracle(Nx, Ny, 2, 2 * np.pi, 2 * np.pi, g=9.81)
    assert np.abs(o.rhs_hos(np.zeros((2, Ny // 2 + 1, Nx), complex))).max() == 0
    u = random_state(Nx, Ny, 2, amp=0.1, seed=8)
    u[1] = 0
    r = o.rhs_hos(u)
    assert np.abs(r[0]).max() == 0
    assert np.abs(r[1] + 9.81 * u[0]).max() < 1e-16


@pytest.mark.parametrize("depth", [0.0, 1.2])
def test_rhs_converges_to_zakharov_with_exact_w(oracle_mod, depth):
    """As M grows, the HOS RHS tends to the exact Zakharov form
    eta_t = -grad eta.grad phi + (1+|grad eta|^2) W,
    phi_t = -g eta - |grad phi|^2/2 + (1+|grad eta|^2) W^2/2
    evaluated with the exact W of a harmonic potential (dealiased alike)."""
    Nx = Ny = 64
    L = 2 * np.pi
    g = 9.81
    u, W, eta, phi = harmonic_state(Nx, Ny, L, depth, amp_eta=0.06, amp_phi=0.2)
    et, pt = zakharov_rhs(eta, phi, W, L, g)
    errs = []
    for nlevs in range(1, 6):
        o = oracle_mod.Oracle(Nx, Ny, nlevs, L, L, g=g, depth=depth)
        r = o.rhs_hos(u)
        ref = np.stack([o.Dealias(to_spec(et)), o.Dealias(to_spec(pt))])
        errs.append(rel_l2(r[0], ref[0]) + rel_l2(r[1], ref[1]))
    for m in range(len(errs) - 1):
        assert errs[m + 1] < 0.5 * errs[m], errs
    assert errs[-1] < 1e-5, errs


@pytest.mark.parametrize("depth", [0.0, 0.8])
def test_truncation_order_in_amplitude(oracle_mod, depth):
    """Scale eta and the harmonic potential by eps.  A sum truncated at order K
    then misses its exact value by O(eps^(K+1)): W^(M), (W^2)^(M) and both RHS
    components (against the Zakharov form with exact W) by eps^(M+1); W^(M-2)
    by eps^(M-1) and (W^2)^(M-2) by eps^(M-1), where those sums are empty
    (M <= 2) by W or W^2 itself, eps^1 and eps^2.  The error slopes between
    eps = 1/4, 1/8 and 1/16 must match: a term dropped, mis-signed or
    mis-weighted at order n <= K leaves a slope of n instead."""
    Nx = Ny = 64
    L = 2 * np.pi
    g = 9.81
    for nlevs in range(0, 5):
        M = nlevs + 1
        o = oracle_mod.Oracle(Nx, Ny, nlevs, L, L, g=g, depth=depth)
        errs = []
        for eps in (0.25, 0.125, 0.0625):
            u, W, eta, phi = harmonic_state(Nx, Ny, L, depth, amp_eta=0.08 * eps, amp_phi=0.3 * eps)
            wM, wM2, w2M, w2M2 = (o.ifft_2d(w) for w in o.Zvel(u))
            et, pt = zakharov_rhs(eta, phi, W, L, g)
            r = o.rhs_hos(u)
            errs.append([np.abs(wM - W).max(), np.abs(wM2 - W).max(), np.abs(w2M - W * W).max(),
                         np.abs(w2M2 - W * W).max(), np.abs(r[0] - o.Dealias(to_spec(et))).max(),
                         np.abs(r[1] - o.Dealias(to_spec(pt))).max()])
        errs = np.array(errs)
        slopes = np.log2(errs[:-1] / errs[1:])
        want = np.array([M + 1, max(M - 1, 1), M + 1, max(M - 1, 2), M + 1, M + 1])
        assert np.all(np.abs(slopes - want) < 0.15), (M, slopes, want)


# ---------------------------------------------------------------- RK (A1, A2)
def test_butcher_order_conditions(oracle_mod):
    o = oracle_mod.Oracle(16, 16, 2, 1.0, 1.0)
    for scheme, order in [(1, 4), (2, 4), (3, 1), (4, 2), (5, 2)]:
        o.lib.or_Setup_TimeScheme(o.ctx, scheme)
        a, b, c = o.tableau()
        assert np.allclose(b.sum(axis=1), a, atol=1e-15)           # row-sum consistency
        assert c.sum() == pytest.approx(1.0, abs=1e-15)
        if order >= 2:
            assert c @ a == pytest.approx(1 / 2, abs=1e-14)
        if order >= 4:
            assert c @ a ** 2 == pytest.approx(1 / 3, abs=1e-14)
            assert c @ b @ a == pytest.approx(1 / 6, abs=1e-14)
            assert c @ a ** 3 == pytest.approx(1 / 4, abs=1e-14)
            assert (c * a) @ b @ a == pytest.approx(1 / 8, abs=1e-14)
            assert c @ b @ a ** 2 == pytest.approx(1 / 12, abs=1e-14)
            assert c @ b @ b @ a == pytest.approx(1 / 24, abs=1e-14)


@pytest.mark.parametrize("scheme", [1, 2, 3, 4, 5])
def test_rk_linear_wave_matches_stability_function(oracle_mod, scheme):
    """A tiny-amplitude single mode evolves as R(dt*A)^n with A = [[0,k],[-g,0]]
    (deep water), R the scheme's stability function (textbook polynomials for
    RK4 / Euler / 2-stage, the tableau formula for Dormand-Prince)."""
    Nx = Ny = 16
    L, g, dt, n = 2 * np.pi, 9.81, 0.05, 40
    o = oracle_mod.Oracle(Nx, Ny, 1, L, L, g=g, scheme=scheme)
    x, y = grid(Nx, Ny, L, L)
    k, a0 = 2.0, 1e-9
    u = np.zeros((2, Ny // 2 + 1, Nx), complex)
    u[0] = to_spec(a0 * np.cos(k * x) + 0 * y)
    o.run(u, 0.0, dt, n)
    A = np.array([[0.0, k], [-g, 0.0]]) * dt
    I2 = np.eye(2)
    if scheme == 2:
        R = I2 + A + A @ A / 2 + A @ A @ A / 6 + A @ A @ A @ A / 24
    elif scheme == 3:
        R = I2 + A
    elif scheme in (4, 5):
        R = I2 + A + A @ A / 2
    else:
        aa, bb, cc = o.tableau()
        S = len(cc)
        # K_s = A (y + sum_p b_sp K_p) -> solve the block system
        Kmat = np.zeros((2 * S, 2))
        for s in range(S):
            acc = I2.copy()
            for p in range(s):
                acc = acc + bb[s, p] * Kmat[2 * p:2 * p + 2]
            Kmat[2 * s:2 * s + 2] = A @ acc
        R = I2 + sum(cc[s] * Kmat[2 * s:2 * s + 2] for s in range(S))
    v = np.linalg.matrix_power(R, n) @ np.array([a0, 0.0])
    j = int(round(k * L / (2 * np.pi)))
    got_eta = 2 * u[0, 0, j].real
    got_phi = 2 * u[1, 0, j].real
    assert got_eta == pytest.approx(v[0], rel=1e-6, abs=1e-18)
    assert got_phi == pytest.approx(v[1], rel=1e-6, abs=1e-18)


def test_stage_time_not_advanced(oracle_mod):
    """time_schemes_par.c:70: every stage sees t_n; with the ramp on, one RK4
    step equals four RHS calls at t_n combined by the Butcher weights."""
    Nx = Ny = 16
    o = oracle_mod.Oracle(Nx, Ny, 2, 2 * np.pi, 2 * np.pi, rampflg=1, Tramp=0.7)
    u = random_state(Nx, Ny, 2, amp=0.05, seed=12)
    t, dt = 0.4, 0.05
    k1 = o.rhs_hos(u, t)
    k2 = o.rhs_hos(u + 0.5 * dt * k1, t)
    k3 = o.rhs_hos(u + 0.5 * dt * k2, t)
    k4 = o.rhs_hos(u + dt * k3, t)
    ref = u + dt * (k1 + 2 * k2 + 2 * k3 + k4) / 6
    got = u.copy()
    t2, _ = o.step(got, t, dt)
    assert t2 == pytest.approx(t + dt)
    assert rel_l2(got, ref) < 1e-14
    assert o.RampFun(t) == pytest.approx(1 - np.exp(-t * t / 0.49))


def _linear_rhs(u, Nx, Ny, nlevs, Lx, Ly, g):
    """Masked linear RHS (W^(1) = |k| phi-hat, -g eta-hat) from numpy's own
    wavenumbers (signed kx; the modes where the reference table differs are
    outside the kept band for nlevs >= 1)."""
    kx = 2 * np.pi / Lx * np.fft.fftfreq(Nx, 1.0 / Nx)[None, :]
    ky = 2 * np.pi / Ly * np.arange(Ny // 2 + 1)[:, None]
    kk = np.sqrt(kx * kx + ky * ky)
    keep = ~dealias_mask(Nx, Ny, nlevs)
    return np.stack([kk * u[1] * keep, -g * u[0] * keep])


@pytest.mark.parametrize("nlevs", [1, 2, 4])
def test_ramp_closed_forms(oracle_mod, nlevs):
    """Dommermuth ramp (hosm_2d_par.c:211-229, RampFun :526-530), pinned by
    what the mathematics fixes: R(0) = 0 leaves exactly the masked linear RHS;
    R(50 Tramp) = 1 in double precision gives the unramped RHS; in between the
    RHS is lin + R(t) (unramped - lin) with R(Tramp) = 1 - 1/e and
    R(Tramp/2) = 1 - e^(-1/4).  A sign slip on the g term, R applied to the
    linear part, or t in place of t^2 fails one of these."""
    Nx, Ny = 32, 16
    Lx, Ly, g, Tr = 2 * np.pi, np.pi, 9.81, 2.5
    u = random_state(Nx, Ny, nlevs, amp=0.08, seed=21 + nlevs, Lx=Lx, Ly=Ly)
    plain = oracle_mod.Oracle(Nx, Ny, nlevs, Lx, Ly, g=g)
    ramped = oracle_mod.Oracle(Nx, Ny, nlevs, Lx, Ly, g=g, rampflg=1, Tramp=Tr)
    lin = _linear_rhs(u, Nx, Ny, nlevs, Lx, Ly, g)
    full = plain.rhs_hos(u)
    assert rel_l2(full, lin) > 1e-3  # the nonlinear part is visible
    # t = 0: exactly the linear part
    r0 = ramped.rhs_hos(u, 0.0)
    assert np.abs(r0 - lin).max() <= 1e-15 * np.abs(lin).max()
    # t >> Tramp: the unramped RHS
    assert rel_l2(ramped.rhs_hos(u, 50 * Tr), full) < 1e-14
    # in between: lin + R (full - lin)
    for t, R in ((Tr, 1 - np.exp(-1.0)), (0.5 * Tr, 1 - np.exp(-0.25)), (1.7 * Tr, 1 - np.exp(-1.7 ** 2))):
        want = lin + R * (full - lin)
        assert rel_l2(ramped.rhs_hos(u, t), want) < 1e-14, t
        assert ramped.RampFun(t) == pytest.approx(R, rel=1e-15)
    # the ramp does not depend on the state's time, only on t (stage time t_n)
    assert ramped.RampFun(0.0) == 0.0


# ---------------------------------------------------------------- Hamiltonian (A10)
def test_hamiltonian_parseval(oracle_mod):
    Nx, Ny = 32, 16
    g = 9.81
    o = oracle_mod.Oracle(Nx, Ny, 2, 2 * np.pi, 2 * np.pi, g=g)
    u = random_state(Nx, Ny, 2, amp=0.1, seed=3)
    et = random_state(Nx, Ny, 2, amp=0.2, seed=4)[0]
    H = o.Hamiltonian(u[0], et, u[1])
    eta, phi, etat = (to_phys(f, Nx, Ny) for f in (u[0], u[1], et))
    assert H == pytest.approx(np.mean(g * eta ** 2 + phi * etat), rel=1e-13)


# ---------------------------------------------------------------- whole step
@pytest.mark.slow
def test_stokes_wave_phase_speed(oracle_mod):
    """Deep-water Stokes wave, ka = 0.1, M = 3: the crest moves at
    c = sqrt(g/k) (1 + (ka)^2/2) to O((ka)^4); linear theory would be off by
    (ka)^2/2 = 5e-3."""
    Nx, Ny = 64, 8
    L, g = 2 * np.pi, 9.81
    k, a = 1.0, 0.1
    x, y = grid(Nx, Ny, L, L)
    om = np.sqrt(g * k) * (1 + 0.5 * (k * a) ** 2)
    th = k * x + 0 * y
    eta = a * np.cos(th) + 0.5 * k * a * a * np.cos(2 * th) + 3 / 8 * k * k * a ** 3 * np.cos(3 * th)
    phi = (om / k) * a * np.exp(k * eta) * np.sin(th)  # third-order Stokes potential at z = eta
    nlevs = 2
    o = oracle_mod.Oracle(Nx, Ny, nlevs, L, L, g=g)
    u = np.zeros((2, Ny // 2 + 1, Nx), complex)
    u[0], u[1] = o.Dealias(to_spec(eta)), o.Dealias(to_spec(phi))
    T = 2 * np.pi / om
    n = 160
    dt = 2 * T / n
    t, _ = o.run(u, 0.0, dt, n)
    # phase of the first harmonic after two periods (mod 2 pi): 0 for the
    # nonlinear speed, 2*(ka)^2/2*2pi ~ 0.063 rad ahead for linear theory
    ph = np.angle(u[0, 0, 1])
    drift = np.angle(np.exp(1j * (ph - 0.0)))
    assert abs(drift) < 0.006, drift


def test_oracle_threads_do_not_change_results(oracle_mod):
    """The oracle's transforms run on OpenMP threads (the multi-core CPU
    baseline); every 1-D transform stays on one thread, so two RK4 steps give
    the same bits on 1 and on 4 threads."""
    Nx, Ny, nlevs = 64, 32, 3
    u = random_state(Nx, Ny, nlevs, amp=0.06, seed=44)
    out = []
    for th in (1, 4):
        oracle_mod.set_threads(th)
        o = oracle_mod.Oracle(Nx, Ny, nlevs, 2 * np.pi, np.pi, depth=0.8)
        v = np.ascontiguousarray(u.copy())
        o.run(v, 0.0, 0.02, 2)
        out.append(v)
    oracle_mod.set_threads(1)
    assert np.array_equal(out[0], out[1])
