"""GPU parity of the NTT, polynomial-matrix products and the device PM-basis
(rr_ntt / rr_poly_mat_mul / rr_pm_basis / rr_rank_from_basis) against the
CPU oracle and the reference's golden vectors."""
import os

import numpy as np
import pytest

import pyoracle as O
from rust_rank_amd import synth

P = 2013265921
pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _eng():
    from rust_rank_amd import engine
    assert engine.device_count() >= 1, "no GPU visible"
    return engine


@pytest.mark.parametrize("logn,count", [(1, 3), (4, 5), (6, 100), (10, 2), (10, 5), (11, 3), (14, 1), (12, 7),
                                        (15, 3), (17, 1), (18, 2)])
def test_ntt_matches_oracle(logn, count):
    eng = _eng()
    print(f"ntt logn={logn} count={count}", flush=True)
    rng = np.random.default_rng(logn)
    a = rng.integers(0, P, (count, 1 << logn)).astype(np.uint32)
    f = eng.ntt(a)
    for q in range(count):
        assert np.array_equal(f[q], O.ntt(a[q]))
    assert np.array_equal(eng.ntt(f, inverse=True), a)


@pytest.mark.parametrize("p,shape,lo,hi", [
    (P, (4, 3, 2, 5, 7), 0, 11),          # schoolbook size
    (P, (8, 8, 4, 40, 70), 20, 70),       # NTT, middle window
    (P, (16, 16, 16, 129, 200), 0, 328),  # NTT, full product
    (P, (5, 7, 21, 40, 70), 0, 109),      # NTT, column tiles with a ragged last tile
    (P, (2, 2, 3, 9000, 9000), 0, 17999),   # four-step NTT (2^15 points), full product
    (P, (2, 3, 2, 9000, 20000), 9000, 20000),  # four-step NTT, middle window
    (97, (4, 4, 2, 30, 50), 10, 60),      # direct products, small prime
])
def test_poly_mat_mul_matches_oracle(p, shape, lo, hi):
    eng = _eng()
    rows, inner, cols, la, lb = shape
    print(f"pmm p={p} {shape} [{lo},{hi})", flush=True)
    rng = np.random.default_rng(sum(shape))
    A = rng.integers(0, p, (rows, inner, la)).astype(np.uint32)
    B = rng.integers(0, p, (inner, cols, lb)).astype(np.uint32)
    got = eng.poly_mat_mul(A, B, lo, hi, p)
    want = O.poly_mat_mul(A, B, lo, hi, p)
    assert np.array_equal(got, want)


def test_pm_basis_matches_reference_golden():
    """Small primes where the reference's own PM_Basis is exact."""
    eng = _eng()
    d = np.load(os.path.join(G, "pm_basis.npz"))
    i = 0
    while f"seq{i}" in d:
        p, b, dd, L = (int(x) for x in d[f"meta{i}"])
        print(f"golden pm_basis case {i}: p={p} b={b} d={dd}", flush=True)
        basis, delta = eng.pm_basis(d[f"seq{i}"], b, dd, p)
        assert np.array_equal(basis, d[f"basis{i}"]), i
        assert np.array_equal(delta, d[f"delta{i}"]), i
        assert eng.rank_from_basis(basis, delta, b, p) == int(d[f"rank{i}"][0]), i
        i += 1


@pytest.mark.parametrize("b,d", [(1, 7), (2, 37), (4, 100), (8, 457), (16, 300), (32, 90), (64, 20), (1, 20000),
                                 (2, 17000)])
def test_pm_basis_matches_oracle_big_p(b, d):
    eng = _eng()
    print(f"pm_basis b={b} d={d}", flush=True)
    rng = np.random.default_rng(b * 1000 + d)
    seq = rng.integers(0, P, (d + 2, b * (b + 1) // 2)).astype(np.uint32)
    basis, delta = eng.pm_basis(seq, b, d, P)
    want_b, want_d = O.pm_basis(seq, b, d, P)
    assert np.array_equal(delta, want_d)
    assert np.array_equal(basis, want_b)


def test_rank_plain_small_end_to_end():
    """2000 x 1800 plain workload: device Krylov (b = 8, L = 458) -> device PM-basis
    (d = L - 1) -> rank; equals dense Gaussian elimination."""
    eng = _eng()
    A = synth.plain_small()
    b = 8
    dr, dc, V = synth.preconditioners(1, A.m, A.n, b, P)
    L = synth.seq_len_plain(min(A.m, A.n), b)
    with eng.Krylov(A, P, b, dr, dc, V) as K:
        seq = K.run(L)
    basis, delta = eng.pm_basis(seq, b, L - 1, P)
    r = eng.rank_from_basis(basis, delta, b, P)
    assert r == O.dense_rank(A.to_dense(P), P)


def test_rank_streamed_equals_sequence_then_basis():
    """rr_rank on a matrix of at least 2^21 nonzeros and an order of at least
    1024 runs the PM-basis beside the Krylov passes (chunked sequence, G's
    leading coefficients handed over as they appear); its rank equals the
    sequence computed in one run followed by the device PM-basis."""
    from rust_rank_amd import engine as eng
    assert eng.device_count() >= 1, "no GPU visible"
    A = synth.graph_complex(seed=7, m=180000, n=8000, k=12)
    assert A.nnz >= 1 << 21
    b = 8
    L = 2 * min(A.m, A.n) // b + 8
    d = L - 1
    assert d >= 1024
    dr, dc, V = synth.preconditioners(11, A.m, A.n, b, P)
    K = eng.Krylov(A, P, b, dr, dc, V)
    seq = K.run(L)
    K.close()
    basis, delta = eng.pm_basis(seq, b, d, P)
    want = eng.rank_from_basis(basis, delta, b, P)
    got, ms = eng.rank(A, P, b, dr, dc, V)
    print(f"m={A.m} n={A.n} nnz={A.nnz} b={b} d={d}: streamed rank {got} in {ms:.0f} ms, two-step {want}", flush=True)
    assert got == want
