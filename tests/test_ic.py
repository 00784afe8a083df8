"""Properties of the seeded JONSWAP initial condition (hos_ic_jonswap, host C):
closed-form facts of the reference recipe (python/JONSWAP.py, DtoF.py,
CreateSpectra.py) rather than a second copy of its formulas."""
import math

import numpy as np
import pytest


def modes(N, L):
    k1 = 2 * math.pi / L
    kx = np.fft.fftfreq(N, 1.0 / N)[None, :] * k1
    ky = np.arange(N // 2 + 1)[:, None] * k1
    return kx + 0 * ky, ky + 0 * kx


@pytest.fixture(scope="module")
def setup(hos):
    N, kp = 128, 1.0
    L = 8 * 2 * math.pi / kp
    return hos, N, L, kp


def test_std_cut_and_zero_lines(setup):
    hos, N, L, kp = setup
    u = hos.jonswap(N, N, L, L, kp=kp, gamma=3.3, eps=0.1, seed=5)
    kx, ky = modes(N, L)
    k = np.hypot(kx, ky)
    # std(eta) = Hs/4 = eps/(2 kp): Parseval over the half spectrum
    var = np.sum(np.abs(u[0, 0]) ** 2) + 2 * np.sum(np.abs(u[0, 1:]) ** 2)
    assert math.sqrt(var) == pytest.approx(0.1 / (2 * kp), rel=1e-13)
    assert np.all(u[:, k > 4 * kp] == 0)
    assert np.all(u[:, 0, 0] == 0)
    assert np.all(u[:, :, N // 2] == 0) and np.all(u[:, N // 2, :] == 0)


def test_linear_right_going_phase(setup):
    """phi-hat = -i sign(kx) (g/omega) eta-hat: every mode is a linear wave
    travelling within 90 degrees of +x (cos^2 spreading, DtoF.py:55-56)."""
    hos, N, L, kp = setup
    for depth in (0.0, 1.3):
        u = hos.jonswap(N, N, L, L, kp=kp, gamma=2.0, eps=0.1, depth=depth, seed=1)
        kx, ky = modes(N, L)
        k = np.hypot(kx, ky)
        om = np.sqrt(9.81 * k * (np.tanh(k * depth) if depth > 0 else 1.0))
        nz = np.abs(u[0]) > 0
        assert nz.sum() > 100
        ratio = u[1][nz] / u[0][nz]
        want = -1j * np.sign(kx[nz]) * 9.81 / om[nz]
        assert np.allclose(ratio, want, rtol=1e-12)


def test_peak_enhancement_and_widths(setup):
    """With the same phases, |eta|^2 for gamma over gamma = 1 is
    gamma^exp(-(w-wp)^2/(2 s^2 wp^2)): exactly gamma at the peak, s = 0.07
    below it and 0.09 above (JONSWAP.py:1-9, 40-49)."""
    hos, N, L, kp = setup
    gamma = 4.0
    a = hos.jonswap(N, N, L, L, kp=kp, gamma=gamma, eps=0.1, seed=2)
    b = hos.jonswap(N, N, L, L, kp=kp, gamma=1.0, eps=0.1, seed=2)
    nz = (np.abs(a[0]) > 0) & (np.abs(b[0]) > 0)
    r = np.abs(a[0][nz]) ** 2 / np.abs(b[0][nz]) ** 2
    kx, ky = modes(N, L)
    w = np.sqrt(9.81 * np.hypot(kx, ky))[nz]
    wp = math.sqrt(9.81 * kp)
    delta = np.log(r / r.min() * 1.0) / math.log(gamma)  # r.min() is the far tail, delta -> 0 there
    # the rescale makes r proportional, so compare shapes through the peak mode
    peak = np.argmin(np.abs(w - wp))
    assert abs(w[peak] - wp) < 1e-12  # kp lies on the grid
    c = r[peak] / gamma  # normalisation common to all modes
    delta = np.log(r / c) / math.log(gamma)
    for side, sig in ((w < wp, 0.07), (w > wp, 0.09)):
        want = np.exp(-(w[side] - wp) ** 2 / (2 * sig ** 2 * wp ** 2))
        assert np.allclose(delta[side], want, atol=1e-10)


def test_cos2_spreading(setup):
    """Two modes of equal |k|: |eta|^2 ratio = cos^2 theta ratio."""
    hos, N, L, kp = setup
    u = hos.jonswap(N, N, L, L, kp=kp, gamma=3.3, eps=0.1, seed=3)
    kx, ky = modes(N, L)
    # (3,4) and (5,0) in grid units both have |k| = 5 dk
    a = np.abs(u[0, 4, 3]) ** 2
    b = np.abs(u[0, 0, 5]) ** 2
    assert a / b == pytest.approx(math.cos(math.atan2(4, 3)) ** 2, rel=1e-12)
    row = u[0, 0]  # ky = 0 row of a real field: Hermitian in kx
    assert np.all(row[1:] == np.conj(row[1:][::-1]))


def test_seed_determinism_and_uniform(hos):
    u1 = hos.jonswap(64, 64, 50.0, 50.0, seed=11)
    u2 = hos.jonswap(64, 64, 50.0, 50.0, seed=11)
    u3 = hos.jonswap(64, 64, 50.0, 50.0, seed=12)
    assert np.array_equal(u1, u2) and not np.array_equal(u1, u3)
    lib = hos.load()
    v = np.array([lib.hos_ic_uniform(7, i, j) for i in range(64) for j in range(64)])
    assert v.min() >= 0 and v.max() < 1 and abs(v.mean() - 0.5) < 0.02
