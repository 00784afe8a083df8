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


def test_uneven_mix_at_its_own_shapes(hos, oracle_mod):
    """The first ten requests of the 96-request uneven batch at their own
    shapes (128^2, 256^2, 512^2, 1024^2 and 2048x512 grids, M = 2..6, kp*h =
    0.8, 1.5 and deep water), with short horizons: every result equals the
    same request run alone, bit for bit; a 256^2 M=5 deep-water request served
    in FP64 matches the oracle."""
    import hos_amd.serve_loop as sl
    print("serve uneven mix at its own shapes", flush=True)
    reqs = []
    for i in range(10):
        r = sl.mix_request(i, hmin=6, hmax=12)
        reqs.append((r, sl.initial_state(r)))
    seen = {(r["Nx"], r["Ny"]) for r, _ in reqs}
    assert seen == set(sl.GRIDS)
    assert {r["nlevs"] + 1 for r, _ in reqs} >= {2, 4, 5, 6}
    assert {r["depth"] for r, _ in reqs} == {0.8, 1.5, 0.0}
    srv = hos.Server(hbm_budget=250e9, precision=hos.HOS_FP32, steps_per_launch=4)
    srv.start()
    ids = [sl.submit(srv, r, u0) for r, u0 in reqs]
    srv.drain(300.0)
    srv.stop()
    st = srv.stats()
    assert st["completed"] == len(reqs)
    for (r, u0), rid in zip(reqs, ids):
        u, info = srv.result(rid)
        assert np.array_equal(u, sl.run_alone(r, u0, hos.HOS_FP32)), (r["Nx"], r["Ny"], r["nlevs"])
    srv.close()
    # FP64 request 10 of the mix: 256^2, M = 5, deep water
    r = sl.mix_request(10, hmin=4, hmax=4)
    assert (r["Nx"], r["Ny"], r["nlevs"], r["depth"]) == (256, 256, 4, 0.0)
    u0 = sl.initial_state(r)
    srv = hos.Server(hbm_budget=50e9, precision=hos.HOS_FP64)
    rid = sl.submit(srv, r, u0)
    srv.drain(120.0)
    u, info = srv.result(rid)
    srv.close()
    o = oracle_mod.Oracle(r["Nx"], r["Ny"], r["nlevs"], r["Lx"], r["Ly"], g=r["g"], depth=r["depth"])
    ref = np.ascontiguousarray(u0.copy())
    t, H = o.run(ref, 0.0, r["dt"], r["nsteps"])
    assert rel_l2(u, ref) < 1e-12
    assert info["H"] == pytest.approx(H[-1], rel=1e-11)


def test_failed_launch_fails_its_requests(hos):
    """A pool launch that fails (forced on the host) marks that pool's
    requests failed, whatever steps they had done, instead of reporting them
    done; the server then serves new requests of the same shape normally."""
    import hos_amd.serve_loop as sl
    print("serve failed launch", flush=True)
    srv = hos.Server(hbm_budget=20e9, precision=hos.HOS_FP32, max_slots=4, steps_per_launch=2)
    reqs = []
    for i in range(3):
        r = sl.mix_request(i, grids=[(64, 64)], orders=(3,), hmin=10, hmax=10)
        reqs.append((r, sl.initial_state(r)))
    ids = [sl.submit(srv, r, u0) for r, u0 in reqs]
    assert srv.step() == 3  # admitted, first 2 steps in flight
    srv.fail_launches(1)
    t_end = time.time() + 30
    failed = False
    while time.time() < t_end and not failed:
        try:
            srv.step()
        except hos.HosError:
            failed = True
        time.sleep(0.001)
    assert failed
    for rid in ids:
        assert srv.status(rid) == hos.REQ_FAILED
        with pytest.raises(hos.HosError):
            srv.result(rid)
    more = []
    for i in range(3, 5):
        r = sl.mix_request(i, grids=[(64, 64)], orders=(3,), hmin=5, hmax=5)
        more.append((r, sl.initial_state(r)))
    ids2 = [sl.submit(srv, r, u0) for r, u0 in more]
    srv.drain(120.0)
    for (r, u0), rid in zip(more, ids2):
        u, _ = srv.result(rid)
        assert np.array_equal(u, sl.run_alone(r, u0, hos.HOS_FP32))
    srv.close()


def test_poisson_arrivals_overlap(hos):
    """Poisson arrivals at 20 requests/s for 3 s into a running server (the
    request mix of the uneven batch on its smaller grids, horizons of 20-60
    steps): requests that arrive while others run are admitted as slots free
    up, pools of different shapes have steps in flight at the same time, and a
    sample of results equals the same requests run alone."""
    import hos_amd.serve_loop as sl
    print("serve poisson arrivals", flush=True)
    rng = np.random.default_rng(3)
    srv = hos.Server(hbm_budget=100e9, precision=hos.HOS_FP32)
    srv.start()
    reqs, ids = [], []
    t0 = time.time()
    t_next = rng.exponential(1 / 20)
    while t_next < 3.0:
        now = time.time() - t0
        if now < t_next:
            time.sleep(t_next - now)
        r = sl.mix_request(len(reqs), grids=[(128, 128), (256, 256), (512, 512)], hmin=20, hmax=60)
        u0 = sl.initial_state(r)
        reqs.append((r, u0))
        ids.append(sl.submit(srv, r, u0))
        t_next += rng.exponential(1 / 20)
    srv.drain(300.0)
    srv.stop()
    st = srv.stats()
    assert st["completed"] == len(reqs) >= 30
    assert st["peak_running"] >= 2 and st["pools"] >= 3
    for k, ((r, u0), rid) in enumerate(zip(reqs, ids)):
        u, info = srv.result(rid)
        assert info["done"] >= info["admit"] >= info["submit"]
        if k % 8 == 0:
            assert np.array_equal(u, sl.run_alone(r, u0, hos.HOS_FP32))
    srv.close()
