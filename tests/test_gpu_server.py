"""GPU: the continuous-batching server returns, for every request, the rank of
the single-matrix path, whatever else is in flight; a bad request fails alone;
admission respects the HBM budget."""
import time

import numpy as np
import pytest

import pyoracle as O
from rust_rank_amd import synth

P = 2013265921
pytestmark = pytest.mark.gpu


def _eng():
    from rust_rank_amd import engine
    assert engine.device_count() >= 1, "no GPU visible"
    return engine


def _requests(n, seed, lo=150, hi=1200):
    out = []
    for i in range(n):
        A = synth.batch_matrix(seed, i, rows_lo=lo, rows_hi=hi)
        b = (4, 8, 16)[i % 3]
        dr, dc, V = synth.preconditioners(seed * 1000 + i, A.m, A.n, b, P)
        out.append((A, b, dr, dc, V))
    return out


def _collect(srv, ids, timeout=120):
    got = {}
    t0 = time.time()
    while len(got) < len(ids) and time.time() - t0 < timeout:
        for r in srv.poll(wait_ms=200):
            got[r["id"]] = r
    assert len(got) == len(ids), f"only {len(got)} of {len(ids)} results"
    return got


def test_server_results_equal_alone_and_in_mix():
    eng = _eng()
    reqs = _requests(10, 31)
    alone = [eng.rank(A, P, b, dr, dc, V)[0] for (A, b, dr, dc, V) in reqs]
    with eng.Server(P, 64 << 30) as srv:
        ids = []
        for k, (A, b, dr, dc, V) in enumerate(reqs):
            ids.append(srv.submit(A, b, dr, dc, V))
            if k % 3 == 0:
                time.sleep(0.01)  # arrivals while others are in flight
        got = _collect(srv, ids)
        print(srv.stats(), flush=True)
    for i, rid in enumerate(ids):
        r = got[rid]
        assert r["status"] == 0 and r["rank"] == alone[i], (i, r, alone[i])
        assert r["latency_ms"] >= r["queue_ms"] >= 0
    A0 = reqs[0][0]
    assert alone[0] == O.dense_rank(A0.to_dense(P), P)


def test_server_isolates_a_failing_request():
    eng = _eng()
    reqs = _requests(4, 32)
    with eng.Server(P, 64 << 30) as srv:
        ids = [srv.submit(A, b, dr, dc, V) for (A, b, dr, dc, V) in reqs[:2]]
        A, b, dr, dc, V = reqs[2]
        bad = dr.copy()
        bad[0] = 0  # preconditioner entries must be in [1, p)
        bad_id = srv.submit(A, b, bad, dc, V)
        ids += [srv.submit(A, b, dr, dc, V) for (A, b, dr, dc, V) in reqs[3:]]
        got = _collect(srv, ids + [bad_id])
    assert got[bad_id]["status"] != 0 and got[bad_id]["rank"] == -1
    for rid, (A, b, dr, dc, V) in zip(ids, reqs[:2] + reqs[3:]):
        assert got[rid]["status"] == 0
        assert got[rid]["rank"] == eng.rank(A, P, b, dr, dc, V)[0]


def test_server_small_budget_serialises_admission():
    eng = _eng()
    reqs = _requests(5, 33)
    with eng.Server(P, 1) as srv:  # budget below any request: they run one at a time
        ids = [srv.submit(A, b, dr, dc, V) for (A, b, dr, dc, V) in reqs]
        got = _collect(srv, ids)
        st = srv.stats()
    assert st["max_active"] == 1
    for rid, (A, b, dr, dc, V) in zip(ids, reqs):
        assert got[rid]["rank"] == eng.rank(A, P, b, dr, dc, V)[0]


def test_server_streamed_pm_basis_equals_single():
    """A request of order >= 1024 starts its PM-basis at admission on the
    leading coefficients of G, among short requests; its rank equals the
    single-matrix path's."""
    eng = _eng()
    A = synth.batch_matrix(6, 3, rows_lo=2600, rows_hi=2700)
    dr, dc, V = synth.preconditioners(9, A.m, A.n, 4, P)
    reqs = [(A, 4, dr, dc, V)] + _requests(4, 37)
    alone = [eng.rank(M, P, b, r, c, v)[0] for (M, b, r, c, v) in reqs]
    with eng.Server(P, 64 << 30) as srv:
        ids = [srv.submit(M, b, r, c, v) for (M, b, r, c, v) in reqs]
        got = _collect(srv, ids)
    print(f"m={A.m} n={A.n} b=4: rank {got[ids[0]]['rank']} alone {alone[0]}", flush=True)
    for i, rid in enumerate(ids):
        assert got[rid]["status"] == 0 and got[rid]["rank"] == alone[i], (i, got[rid], alone[i])
