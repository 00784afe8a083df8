"""Host logic of hos_amd.serve_loop (no GPU): the request mix and the error measure."""
import numpy as np

from hos_amd import serve_loop as sl


def test_eta_error_uses_eta_only():
    rng = np.random.default_rng(0)
    ref = rng.standard_normal((2, 5, 8)) + 1j * rng.standard_normal((2, 5, 8))
    u = ref.copy()
    u[1] += 10.0  # phi differs, eta does not
    assert sl.eta_rel_err(u, ref) == 0.0
    u[0] += 0.01 * ref[0]
    assert abs(sl.eta_rel_err(u, ref) - 0.01) < 1e-12


def test_mix_request_is_seeded_and_in_range():
    a, b = sl.mix_request(7), sl.mix_request(7)
    assert a == b
    for i in range(50):
        r = sl.mix_request(i, hmin=200, hmax=2000)
        assert (r["Nx"], r["Ny"]) in sl.GRIDS
        assert 2 <= r["nlevs"] + 1 <= 6
        assert 200 <= r["nsteps"] <= 2000
        assert 0.04 <= r["eps"] <= 0.13 and 1.0 <= r["gamma"] <= 7.0
