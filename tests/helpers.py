"""Shared input builders for the tests (seeded, reference layouts)."""
import numpy as np


def grid(Nx, Ny, Lx, Ly):
    x = np.arange(Nx)[:, None] * Lx / Nx
    y = np.arange(Ny)[None, :] * Ly / Ny
    return x, y


def dealias_mask(Nx, Ny, nlevs):
    """True where operators_par.c:314 zeroes a mode (independent restatement)."""
    mx = int(np.floor(0.5 * Nx / (1 + 0.5 * nlevs)))
    my = int(np.floor(0.5 * Ny / (1 + 0.5 * nlevs)))
    i = np.arange(Nx)[None, :]
    j = np.arange(Ny // 2 + 1)[:, None]
    return ((i >= mx) & (i < Nx - mx + 1)) | (j >= my)


def random_state(Nx, Ny, nlevs, amp=0.05, slope=2.0, seed=0, Lx=2 * np.pi, Ly=2 * np.pi):
    """A masked, Hermitian-consistent random state [2][Ny/2+1][Nx] built by
    r2c (numpy) of smooth random real fields of rms ~amp."""
    rng = np.random.default_rng(seed)
    out = np.zeros((2, Ny // 2 + 1, Nx), complex)
    kx = np.fft.fftfreq(Nx, 1.0 / Nx)[:, None]
    ky = np.fft.rfftfreq(Ny, 1.0 / Ny)[None, :]
    kk = np.sqrt(kx ** 2 + ky ** 2)
    for c in range(2):
        f = rng.standard_normal((Nx, Ny))
        F = np.fft.rfft2(f) / (1.0 + kk) ** slope
        f = np.fft.irfft2(F, s=(Nx, Ny))
        f *= amp / f.std()
        out[c] = (np.fft.rfft2(f) / (Nx * Ny)).T
    out[:, dealias_mask(Nx, Ny, nlevs)] = 0
    return out


def to_phys(hu, Nx, Ny):
    """reference spectral [Ny/2+1][Nx] (normalised) -> real [Nx][Ny] via numpy."""
    return np.fft.irfft2(np.asarray(hu).T * (Nx * Ny), s=(Nx, Ny))


def to_spec(u):
    Nx, Ny = u.shape
    return (np.fft.rfft2(u) / (Nx * Ny)).T


def harmonic_state(Nx, Ny, L, depth, amp_eta=0.08, amp_phi=0.3):
    """eta (smooth, two modes) and phi_s = Phi(x, y, eta) for a potential Phi
    that solves Laplace's equation with a flat bed at -depth (deep if 0).
    Returns the spectral state and the exact W = dPhi/dz at z = eta."""
    x, y = grid(Nx, Ny, L, L)
    k0 = 2 * np.pi / L
    eta = amp_eta * (np.cos(k0 * x) + 0.5 * np.sin(k0 * (x + 2 * y)))
    kx, ky = 2 * k0, k0
    k = np.hypot(kx, ky)
    th = kx * x + ky * y
    if depth > 0:
        Z = np.cosh(k * (eta + depth)) / np.cosh(k * depth)
        dZ = k * np.sinh(k * (eta + depth)) / np.cosh(k * depth)
    else:
        Z = np.exp(k * eta)
        dZ = k * np.exp(k * eta)
    phi = amp_phi * Z * np.cos(th)
    W = amp_phi * dZ * np.cos(th)
    u = np.zeros((2, Ny // 2 + 1, Nx), complex)
    u[0], u[1] = to_spec(eta), to_spec(phi)
    return u, W, eta, phi


def zakharov_rhs(eta, phi, W, L, g):
    """The exact Zakharov right-hand sides on a square periodic grid of side L,
    with spectral gradients of eta and phi_s and the exact W of the potential."""
    k0 = np.fft.fftfreq(eta.shape[0], L / eta.shape[0] / (2 * np.pi))

    def d(f, ax):
        kk = k0[:, None] if ax == 0 else k0[None, :]
        return np.real(np.fft.ifft(1j * kk * np.fft.fft(f, axis=ax), axis=ax))

    ex, ey, px, py = d(eta, 0), d(eta, 1), d(phi, 0), d(phi, 1)
    G = ex ** 2 + ey ** 2
    et = -(ex * px + ey * py) + (1 + G) * W
    pt = -g * eta - 0.5 * (px ** 2 + py ** 2) + 0.5 * (1 + G) * W * W
    return et, pt
