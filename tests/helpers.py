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
