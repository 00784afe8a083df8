"""End to end through the reference-compatible boundary: hos_mkic writes a
reference-format input pair, hos_driver (the main_2d_par.c call sequence
over libhos_ref.so) runs it and writes reference-format snapshots, which must
match the oracle run from the same input."""
import math
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import rel_l2
from h5read import have_h5dump, read

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "hos_amd")


@pytest.mark.skipif(not have_h5dump(), reason="h5dump not in this image")
@pytest.mark.parametrize("scheme,nlevs", [(2, 2), (4, 1)])
def test_driver_matches_oracle(tmp_path, oracle_mod, scheme, nlevs):
    print(f"driver scheme={scheme} nlevs={nlevs}", flush=True)
    Nx = Ny = 64
    kp = 1.0
    L = 4 * 2 * math.pi / kp
    Tp = 2 * math.pi / math.sqrt(9.81 * kp)
    dt = Tp / 32
    nsteps = 12
    T = 100  # longer than the run; the run stops after nsteps (-N)
    d = str(tmp_path)
    subprocess.run([os.path.join(BIN, "hos_mkic"), str(Nx), str(Ny), repr(L), repr(L), "9.81", "0", "1", "3.3",
                    "0.1", "3", str(T), repr(dt * 4), "1", d], check=True, timeout=120)
    out = subprocess.run([os.path.join(BIN, "hos_driver"), "-d", d, "-n", str(nlevs), "-s", str(scheme), "-t",
                          repr(dt), "-N", str(nsteps)], check=True, capture_output=True, text=True, timeout=120)
    print(out.stdout[-400:], flush=True)
    files = sorted(f for f in os.listdir(d) if re.fullmatch(r"data\d+\.1\.h5", f))
    assert len(files) == 1 + nsteps // 4  # data0 + one per dtsave crossing
    last = os.path.join(d, f"data{nsteps // 4}.1.h5")
    assert read(last, "/time")[0] == pytest.approx(nsteps * dt, rel=1e-12)
    assert read(last, "/Nx")[0] == Nx and read(last, "/g")[0] == 9.81
    eta0 = read(os.path.join(d, "initdata.1.h5"), "/eta0", (Nx, Ny))
    phi0 = read(os.path.join(d, "initdata.1.h5"), "/phi0", (Nx, Ny))
    # data0 holds the raw initial condition (main_2d_par.c:271-273)
    assert np.array_equal(read(os.path.join(d, "data0.1.h5"), "/eta", (Nx, Ny)), eta0)

    o = oracle_mod.Oracle(Nx, Ny, nlevs, L, L, g=9.81, scheme=scheme)
    u = np.zeros((2, Ny // 2 + 1, Nx), complex)
    u[0] = o.Dealias(o.fft_2d(eta0))
    u[1] = o.Dealias(o.fft_2d(phi0))
    t, H = o.run(u, 0.0, dt, nsteps)
    eta = read(last, "/eta", (Nx, Ny))
    phi = read(last, "/phi", (Nx, Ny))
    assert rel_l2(eta, o.ifft_2d(u[0])) < 1e-12
    assert rel_l2(phi, o.ifft_2d(u[1])) < 1e-12
    # runtime.dat: "t H" per step as time_schemes_par.c:88
    rt = np.loadtxt(os.path.join(d, "runtime.dat"))
    assert rt.shape == (nsteps, 2)
    assert np.allclose(rt[:, 0], np.arange(nsteps) * dt, atol=1e-9)
    assert np.allclose(rt[:, 1], H, rtol=1e-12)
    # XDMF index: one grid per snapshot over physical coordinates
    xmf = open(os.path.join(d, "HOS_data.xmf")).read()
    assert xmf.count("<Grid Name=\"Cells\"") == 1 + nsteps // 4 and xmf.rstrip().endswith("</Xdmf>")
    xy = read(os.path.join(d, "data_coordinates.h5"), "/xy", (Nx * Ny, 2))
    assert xy[Ny + 1, 0] == pytest.approx(L / Nx) and xy[1, 1] == pytest.approx(L / Ny)


@pytest.mark.skipif(not have_h5dump(), reason="h5dump not in this image")
def test_driver_ascii_parameters(tmp_path):
    """initpars.txt (key = value) in place of initpars.h5 gives the same run."""
    print("driver ascii", flush=True)
    Nx = Ny = 32
    L = 2 * 2 * math.pi
    dt = 0.05
    runs = []
    for form in ("h5", "txt"):
        d = tmp_path / form
        d.mkdir()
        subprocess.run([os.path.join(BIN, "hos_mkic"), str(Nx), str(Ny), repr(L), repr(L), "9.81", "0", "1", "3.3",
                        "0.1", "4", "100", "0.2", "1", str(d)], check=True, timeout=120)
        args = ["-n", "2", "-s", "2", "-t", repr(dt)]
        if form == "txt":
            os.remove(d / "initpars.h5")
            (d / "initpars.txt").write_text(
                f"# same run as initpars.h5\nNx = {Nx}\nNy = {Ny}\nLx = {L!r}\nLy = {L!r}\ng = 9.81\n"
                "runsubid = 1\nT = 100\ndtsave = 0.2\nsaveflg = 1\nrampflg = 0\nTramp = 1\n"
                f"nlevs = 2\nscheme = 2\ndt = {dt!r}\n")
            args = []
        subprocess.run([os.path.join(BIN, "hos_driver"), "-d", str(d), *args, "-N", "8"], check=True,
                       capture_output=True, timeout=120)
        runs.append(read(str(d / "data2.1.h5"), "/eta"))
    assert np.array_equal(runs[0], runs[1])
