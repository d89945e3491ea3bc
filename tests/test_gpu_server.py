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


def test_server_degenerate_shapes():
    """Degenerate requests (no nonzeros, 1 x 1, one row, one column, b above
    n) among ordinary ones: each served rank equals dense elimination."""
    eng = _eng()
    reqs = [(synth.csr_from_dense(D), 2) for D in synth.edge_shapes(P).values()]
    reqs += [(A, b) for A, b, *_ in _requests(3, 29)]
    srv = eng.Server(P, 8 << 30)
    ids = []
    for i, (A, b) in enumerate(reqs):
        dr, dc, V = synth.preconditioners(800 + i, A.m, A.n, b, P)
        ids.append(srv.submit(A, b, dr, dc, V))
    got = _collect(srv, ids)
    srv.close()
    for i, (A, b) in enumerate(reqs):
        r = got[ids[i]]
        assert r["status"] == 0 and r["rank"] == O.dense_rank(A.to_dense(P), P), (i, r)


def test_server_other_prime_matches_oracle():
    """The served form mod p = 3323: the PM-basis rounds of the requests run
    their products through the three CRT primes; every rank equals the
    oracle's pipeline."""
    from gpu_cases import p3323_cases
    eng = _eng()
    p, cases = p3323_cases()
    with eng.Server(p, 8 << 30) as srv:
        ids = [srv.submit(A, b, dr, dc, V, L) for (A, b, dr, dc, V, L, _) in cases]
        got = _collect(srv, ids)
    for i, c in enumerate(cases):
        r = got[ids[i]]
        assert r["status"] == 0 and r["rank"] == c[6], (i, r, c[6])


def test_server_lane_width_64():
    """b = 64 and b = 48 requests beside b = 8 ones: every served rank equals
    dense elimination."""
    eng = _eng()
    reqs = []
    for i, b in enumerate((64, 8, 48, 64, 8)):
        A = synth.batch_matrix(73, i, rows_lo=300, rows_hi=900)
        reqs.append((A, b) + tuple(synth.preconditioners(970 + i, A.m, A.n, b, P)))
    with eng.Server(P, 8 << 30) as srv:
        ids = [srv.submit(A, b, dr, dc, V) for (A, b, dr, dc, V) in reqs]
        got = _collect(srv, ids)
    for i, (A, b, *_rest) in enumerate(reqs):
        r = got[ids[i]]
        assert r["status"] == 0 and r["rank"] == O.dense_rank(A.to_dense(P), P), (i, b, r)


def test_server_general_values():
    """Requests with values uniform in [1, p) at lane widths 8, 32 and 64:
    every served rank equals dense elimination."""
    eng = _eng()
    reqs = []
    for i, (b, m, n) in enumerate(((8, 400, 370), (32, 500, 450), (64, 450, 500), (8, 300, 280))):
        A = synth.uniform_sparse(77, i, m, n, 3, 9, P, small=False)
        reqs.append((A, b) + tuple(synth.preconditioners(1010 + i, A.m, A.n, b, P)))
    with eng.Server(P, 8 << 30) as srv:
        ids = [srv.submit(A, b, dr, dc, V) for (A, b, dr, dc, V) in reqs]
        got = _collect(srv, ids)
    for i, (A, b, *_rest) in enumerate(reqs):
        r = got[ids[i]]
        assert r["status"] == 0 and r["rank"] == O.dense_rank(A.to_dense(P), P), (i, b, r)


def test_server_split_rows_match_oracle():
    """Requests whose A^T rows are cut into segments: served ranks equal the
    oracle's pipeline."""
    from gpu_cases import split_row_cases
    eng = _eng()
    cases = split_row_cases()
    with eng.Server(P, 8 << 30) as srv:
        ids = [srv.submit(A, b, dr, dc, V, L) for (A, b, dr, dc, V, L, _) in cases]
        got = _collect(srv, ids)
    for i, c in enumerate(cases):
        r = got[ids[i]]
        assert r["status"] == 0 and r["rank"] == c[6], (i, r, c[6])


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


@pytest.mark.parametrize("stage", ["tick", "pm"])
def test_server_request_failing_after_admission(stage, monkeypatch):
    """A request admitted and set up fails in a Krylov tick or in a PM-basis
    round (failure forced from the host before its work is queued): it is
    reported with its own status (RR_ENOMEM) and every other request, ticked
    and reduced in the same launches, still gets the single-path rank."""
    eng = _eng()
    reqs = _requests(6, 34)
    alone = [eng.rank(A, P, b, dr, dc, V)[0] for (A, b, dr, dc, V) in reqs]
    monkeypatch.setenv("RR_SERVER_FAIL_AT", f"{stage}:3")  # the third request submitted
    with eng.Server(P, 64 << 30) as srv:
        ids = [srv.submit(A, b, dr, dc, V) for (A, b, dr, dc, V) in reqs]
        got = _collect(srv, ids)
    assert ids[2] == 3
    assert got[3]["status"] == -2 and got[3]["rank"] == -1, got[3]
    for i, rid in enumerate(ids):
        if rid != 3:
            assert got[rid]["status"] == 0 and got[rid]["rank"] == alone[i], (i, got[rid], alone[i])


@pytest.mark.timeout(600)
def test_server_served_workload_ranks_match_oracle():
    """Requests of the served workload (8-20 nonzeros per row, +-1/+-2, planted
    deficits, block width from the autotuner) arriving while others are in
    flight: the 16 from the low end of its row range match dense Gaussian
    elimination; one of 17185 rows matches the oracle's whole pipeline
    (Krylov sequence, PM-basis, rank formula) on the same inputs."""
    eng = _eng()
    small = [synth.served_matrix(51, i, 2500) for i in range(16)]
    big = synth.served_matrix(5, 9, 100000)
    assert big.m == 17185
    mats = small + [big]
    bs, _, _ = eng.autotune_b(mats, P)
    reqs = [(A, int(b)) + tuple(synth.preconditioners(600 + i, A.m, A.n, int(b), P))
            for i, (A, b) in enumerate(zip(mats, bs))]
    with eng.Server(P, 64 << 30) as srv:
        ids = []
        for k, (A, b, dr, dc, V) in enumerate(reqs):
            ids.append(srv.submit(A, b, dr, dc, V))
            time.sleep(0.005 * (k % 3))
        got = _collect(srv, ids, timeout=300)
    for i, (A, b, dr, dc, V) in enumerate(reqs[:-1]):
        assert got[ids[i]]["status"] == 0 and got[ids[i]]["rank"] == O.dense_rank(A.to_dense(P), P), i
    A, b, dr, dc, V = reqs[-1]
    L = 2 * min(A.m, A.n) // b + 8
    print(f"oracle pipeline m={A.m} n={A.n} b={b} L={L}", flush=True)
    seq = O.krylov(A, P, b, dr, dc, V, 0, L)
    M, delta = O.pm_basis(seq, b, L - 1, P)
    assert got[ids[-1]]["status"] == 0 and got[ids[-1]]["rank"] == O.rank_from_basis(M, delta, b, P)


def test_server_refuses_orders_beyond_its_transforms():
    """The server keeps twiddle tables for transforms of up to 2^20 points:
    an order of 2^20 or more is refused at submission (rr_rank takes it),
    and the server goes on serving."""
    eng = _eng()
    (A, b, dr, dc, V), = _requests(1, 77)
    with eng.Server(P, 8 << 30) as srv:
        with pytest.raises(eng.RankEngineError, match="order"):
            srv.submit(A, b, dr, dc, V, seq_len=(1 << 20) + 8, order=1 << 20)
        rid = srv.submit(A, b, dr, dc, V)
        got = _collect(srv, [rid])
    assert got[rid]["status"] == 0
    want, _ = eng.rank(A, P, b, dr, dc, V)
    assert got[rid]["rank"] == want
