"""The serving loop (hos_server): continuous batching of uneven requests.
Every result must be bit-identical to the same request run alone, requests
that arrive while others run are admitted as slots free up, and FP64 results
match the oracle."""
import math
import threading
import time

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

GRIDS = [(32, 32), (64, 64), (64, 32), (128, 64)]


def small_mix(n, seed=0):
    import hos_amd.serve_loop as sl
    out = []
    for i in range(n):
        r = sl.mix_request(i + seed, grids=GRIDS, orders=(2, 3, 4), hmin=3, hmax=25)
        out.append((r, sl.initial_state(r)))
    return out


def test_mix_equals_alone_fp32(hos):
    import hos_amd.serve_loop as sl
    print("serve mix fp32", flush=True)
    reqs = small_mix(24)
    srv = hos.Server(hbm_budget=20e9, precision=hos.HOS_FP32, max_slots=4)
    srv.start()
    ids = []
    for k, (r, u0) in enumerate(reqs):  # arrivals spread over the run
        ids.append(sl.submit(srv, r, u0))
        if k % 5 == 4:
            time.sleep(0.02)
    srv.drain(300.0)
    srv.stop()
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["peak_running"] >= 2
    for (r, u0), rid in zip(reqs, ids):
        u, info = srv.result(rid)
        alone = sl.run_alone(r, u0, hos.HOS_FP32)
        assert np.array_equal(u, alone), r
        assert info["done"] >= info["admit"] >= info["submit"]
    srv.close()


def test_admission_by_footprint_and_slots(hos):
    """A budget that holds one small pool: requests wait for memory, then all
    finish; a request larger than the whole budget is rejected at submit."""
    import hos_amd.serve_loop as sl
    print("serve admission", flush=True)
    p = hos.params(Nx=64, Ny=64, nlevs=2, precision=hos.HOS_FP32)
    one = hos.footprint(p, 2)
    srv = hos.Server(hbm_budget=one * 1.5, precision=hos.HOS_FP32, max_slots=2)
    reqs = []
    for i in range(5):
        r = sl.mix_request(i, grids=[(64, 64)], orders=(3,), hmin=4, hmax=4)
        reqs.append((r, sl.initial_state(r)))
    ids = [sl.submit(srv, r, u0) for r, u0 in reqs]
    n = srv.step()
    assert n == 2  # two slots only
    assert srv.status(ids[4]) == hos.REQ_QUEUED
    srv.drain(120.0)
    for (r, u0), rid in zip(reqs, ids):
        u, _ = srv.result(rid)
        assert np.array_equal(u, sl.run_alone(r, u0, hos.HOS_FP32))
    big = sl.mix_request(0, grids=[(1024, 1024)], orders=(5,), hmin=1, hmax=1)
    with pytest.raises(hos.HosError):
        srv.submit(np.zeros((2, 513, 1024), complex), 1024, 1024, 4, big["Lx"], big["Ly"], big["dt"], 1)
    srv.close()


def test_fp64_server_matches_oracle(hos, oracle_mod):
    import hos_amd.serve_loop as sl
    print("serve fp64 vs oracle", flush=True)
    reqs = small_mix(4, seed=100)
    srv = hos.Server(hbm_budget=20e9, precision=hos.HOS_FP64)
    ids = [sl.submit(srv, r, u0) for r, u0 in reqs]
    srv.drain(300.0)
    for (r, u0), rid in zip(reqs, ids):
        u, info = srv.result(rid)
        o = oracle_mod.Oracle(r["Nx"], r["Ny"], r["nlevs"], r["Lx"], r["Ly"], g=r["g"], depth=r["depth"])
        ref = np.ascontiguousarray(u0.copy())
        t, H = o.run(ref, 0.0, r["dt"], r["nsteps"])
        assert rel_l2(u, ref) < 1e-12
        assert info["H"] == pytest.approx(H[-1], rel=1e-11)
    srv.close()
