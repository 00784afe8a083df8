"""The C driver's HDF5 ingest (get_params / get_ic_2d formats) on CPU: no GPU
work happens in -info mode."""
import os
import subprocess

import pytest

from h5read import have_h5dump, read

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "hos_amd", "hos_driver")
REF_FIXTURE = "/root/reference/Csource/2dpar"


def info(d):
    out = subprocess.run([DRIVER, "-d", d, "-info"], check=True, capture_output=True, text=True).stdout
    line = out.strip().splitlines()[-1]
    return {k: float(v) for k, v in (kv.split("=") for kv in line.split())}


@pytest.mark.skipif(not os.path.isdir(REF_FIXTURE), reason="reference checkout not present")
def test_reads_reference_fixture():
    """The reference's own input pair (initpars.h5: int64/float64 scalars,
    initdata.1.h5: float64 [256][256])."""
    assert os.path.exists(DRIVER), "build hos_amd first"
    before = set(os.listdir(REF_FIXTURE))
    v = info(REF_FIXTURE)
    assert set(os.listdir(REF_FIXTURE)) == before  # -info writes nothing
    assert (v["Nx"], v["Ny"], v["runsubid"], v["rampflg"], v["saveflg"]) == (256, 256, 1, 1, 1)
    assert v["Lx"] == pytest.approx(3807.991095260355)
    assert v["g"] == pytest.approx(9.81)
    assert (v["T"], v["dtsave"], v["Tramp"]) == (50, 1, 200)
    assert v["bat_abs"] == 0
    if have_h5dump():
        eta = read(os.path.join(REF_FIXTURE, "initdata.1.h5"), "/eta0", (256, 256))
        assert v["eta_std"] == pytest.approx(eta.std(), rel=1e-12)
        assert v["eta_mean"] == pytest.approx(eta.mean(), abs=1e-15)
