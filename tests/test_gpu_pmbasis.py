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
                                        (15, 3), (17, 1), (18, 2), (21, 1), (22, 2)])
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


def _seq_to_g(seq, b, D):
    """G = [S(x); I_b] as [2b][b][D] (src/sigma_basis.rs:9-54)."""
    G = np.zeros((2 * b, b, D), np.uint32)
    iu = np.triu_indices(b)
    for k in range(D):
        S = np.zeros((b, b), np.uint32)
        S[iu] = seq[k]
        G[:b, :, k] = S + np.triu(S, 1).T
    G[b:, :, 0] = np.eye(b, dtype=np.uint32)
    return G


def _left_combine(u, M, p):
    """sum_i u_i M[i] mod p, exact in uint64 (u split in 16-bit halves)."""
    hi = np.zeros(M.shape[1:], np.uint64)
    lo = np.zeros(M.shape[1:], np.uint64)
    for i in range(M.shape[0]):
        Mi = M[i].astype(np.uint64)
        hi += np.uint64(int(u[i]) >> 16) * Mi
        lo += np.uint64(int(u[i]) & 0xFFFF) * Mi
        if i % 64 == 63:
            hi %= np.uint64(p)
            lo %= np.uint64(p)
    return ((hi % np.uint64(p)) * np.uint64(65536) + lo) % np.uint64(p)


# Primes other than the NTT prime: transform products run mod three NTT primes
# (2013265921, 469762049, 1811939329) and are recombined by Garner's method;
# above a transform prime the operands are reduced first (2^31 - 1 is above
# all three, 10^9 + 7 above the smallest, 3323 below all).
OTHER_P = [2147483647, 1000000007, 3323]


@pytest.mark.parametrize("p", OTHER_P)
@pytest.mark.parametrize("shape,lo,hi", [
    ((8, 8, 4, 40, 70), 20, 70),          # middle window
    ((5, 7, 21, 40, 70), 0, 109),         # column tiles with a ragged last tile
    ((2, 2, 3, 4000, 4000), 0, 7999),     # one-workgroup transforms (2^13 points)
    ((2, 2, 2, 9000, 9000), 0, 17999),    # four-step transforms (2^15 points)
    ((2, 3, 2, 3000, 9000), 3000, 9000),  # middle window, operands of unequal length
])
def test_poly_mat_mul_other_primes_match_oracle(p, shape, lo, hi):
    eng = _eng()
    rows, inner, cols, la, lb = shape
    print(f"pmm p={p} {shape} [{lo},{hi})", flush=True)
    rng = np.random.default_rng(sum(shape) + p % 1000)
    A = rng.integers(0, p, (rows, inner, la)).astype(np.uint32)
    B = rng.integers(0, p, (inner, cols, lb)).astype(np.uint32)
    got = eng.poly_mat_mul(A, B, lo, hi, p)
    want = O.poly_mat_mul(A, B, lo, hi, p)
    assert np.array_equal(got, want)


@pytest.mark.timeout(300)
def test_poly_mat_mul_other_prime_largest_coefficients():
    """The largest exact coefficients the three primes have to hold: inner 128,
    2^19 coefficients per operand, every entry p - 1 with p = 2^31 - 1, so
    coefficient k is 128 c_k (p - 1)^2 (about 2^88) with c_k the number of
    index pairs summing to k; mod p that is 128 c_k, since (p - 1)^2 = 1."""
    eng = _eng()
    p, inner, la = 2147483647, 128, 1 << 19
    print(f"pmm p={p} inner={inner} la=lb={la} all entries p-1", flush=True)
    A = np.full((1, inner, la), p - 1, np.uint32)
    got = eng.poly_mat_mul(A, A.reshape(inner, 1, la), 0, 2 * la - 1, p)
    k = np.arange(2 * la - 1, dtype=np.int64)
    want = (inner * (np.minimum(k, 2 * la - 2 - k) + 1)) % p
    assert np.array_equal(got[0, 0].astype(np.int64), want)


@pytest.mark.parametrize("p,b,d", [(2147483647, 2, 700), (2147483647, 8, 160), (1000000007, 1, 6000),
                                   (3323, 4, 500), (65537, 16, 90)])
def test_pm_basis_other_primes_match_oracle(p, b, d):
    eng = _eng()
    print(f"pm_basis p={p} b={b} d={d}", flush=True)
    rng = np.random.default_rng(b * 1000 + d)
    seq = rng.integers(0, p, (d + 2, b * (b + 1) // 2)).astype(np.uint32)
    basis, delta = eng.pm_basis(seq, b, d, p)
    want_b, want_d = O.pm_basis(seq, b, d, p)
    assert np.array_equal(delta, want_d)
    assert np.array_equal(basis, want_b)


def test_rank_other_prime_end_to_end():
    """rr_rank mod 2^31 - 1 at an order whose PM-basis products run through the
    three primes equals dense elimination."""
    eng = _eng()
    p, b = 2147483647, 4
    A = synth.batch_matrix(31, 0, p=p, rows_lo=1500, rows_hi=1600)
    dr, dc, V = synth.preconditioners(12, A.m, A.n, b, p)
    got, ms = eng.rank(A, p, b, dr, dc, V)
    want = O.dense_rank(A.to_dense(p), p)
    print(f"m={A.m} n={A.n} p={p} b={b}: rank {got} in {ms:.0f} ms, elimination {want}", flush=True)
    assert got == want


@pytest.mark.timeout(900)
def test_pm_basis_b64_four_step_invariants():
    """b = 64, d = 2^15 (2b x 2b = 128 x 128 bases; the top products are
    four-step NTTs at 2^16 points).  The CPU oracle is too slow at this size,
    so the basis is checked against what fixes it: u^T M(x) G(x) v = 0 mod x^d
    for random u, v (products by the oracle's NTT), det M(z) = +-z^(-sum delta)
    at a random z, and every row degree bounded by -delta."""
    eng = _eng()
    b, d = 64, 1 << 15
    print(f"pm_basis b={b} d={d}", flush=True)
    rng = np.random.default_rng(64015)
    seq = rng.integers(0, P, (d + 2, b * (b + 1) // 2)).astype(np.uint32)
    M, delta = eng.pm_basis(seq, b, d, P)
    D = M.shape[2]
    assert M.shape == (2 * b, 2 * b, d + 1)
    # order: u^T M G v vanishes below x^d
    u = rng.integers(0, P, 2 * b)
    v = rng.integers(0, P, b)
    a = _left_combine(u, M, P)                                  # [2b][D]
    bv = _left_combine(v, np.transpose(_seq_to_g(seq, b, D), (1, 0, 2)), P)  # [2b][D]
    N = 1 << (2 * D - 1).bit_length()
    acc = np.zeros(N, np.uint64)
    for j in range(2 * b):
        fa = np.zeros(N, np.uint32)
        fb = np.zeros(N, np.uint32)
        fa[:D] = a[j]
        fb[:D] = bv[j]
        acc = (acc + O.ntt(fa).astype(np.uint64) * O.ntt(fb).astype(np.uint64)) % np.uint64(P)
    c = O.ntt(acc.astype(np.uint32), invert=True)
    assert not c[:d].any(), np.nonzero(c[:d])[0][:5]
    # det M(z) = +-z^(-sum delta)
    z = int(rng.integers(2, P))
    Mz = np.zeros((2 * b, 2 * b), np.uint64)
    for k in range(D - 1, -1, -1):
        Mz = (Mz * np.uint64(z) + M[:, :, k]) % np.uint64(P)
    from test_oracle import _det_mod
    det = _det_mod(Mz, P)
    e = int(-delta.sum())
    assert det in (pow(z, e, P), (-pow(z, e, P)) % P)
    nzmax = np.where(M.any(axis=1), np.arange(D)[None, :], -1).max(axis=1)
    assert np.all(nzmax <= -delta)


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


@pytest.mark.parametrize("p, seed, pseed, rows", [(P, 33, 13, 600), (3323, 35, 14, 500)])
def test_rank_from_wdm_file(tmp_path, p, seed, pseed, rows):
    """A .wdm file holding the oracle's sequence (src/main.rs:397-436: a file
    name ending in .wdm is only analysed): rr_rank_from_wdm at order L - 1
    equals dense elimination; under the reference's order policy (largest
    power of two <= L, src/main.rs:599-606) it equals the oracle's rank from
    its PM-basis of that order; bad orders are refused.  At p = 3323 (the
    reference's default prime) the top products run through the three CRT
    primes, and the power-of-two order is too small for this matrix."""
    eng = _eng()
    A = synth.batch_matrix(seed, 0, p=p, rows_lo=rows, rows_hi=rows + 100)
    b = 4
    dr, dc, V = synth.preconditioners(pseed, A.m, A.n, b, p)
    L = synth.seq_len_plain(min(A.m, A.n), b)
    seq, x = O.krylov(A, p, b, dr, dc, V, 0, L, want_x=True)
    path = str(tmp_path / "s.wdm")
    eng.wdm_save(path, A.m, A.n, p, dr, dc, V, x, seq)
    d2 = 1 << (L.bit_length() - 1)
    assert d2 < L, L
    want2 = O.rank_from_basis(*O.pm_basis(seq, b, d2, p), b, p)
    dense = O.dense_rank(A.to_dense(p), p)
    print(f"p={p} m={A.m} n={A.n} b={b} L={L} power-of-two order {d2}: oracle {want2}, dense {dense}", flush=True)
    assert eng.rank_from_wdm(path) == dense
    assert eng.rank_from_wdm(path, reference_policy=True) == want2
    assert eng.rank_from_wdm(path, order=d2) == want2
    for bad in (L, L + 5):
        with pytest.raises(eng.RankEngineError, match="order"):
            eng.rank_from_wdm(path, order=bad)
    with pytest.raises(eng.RankEngineError):
        eng.rank_from_wdm(str(tmp_path / "missing.wdm"))


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


# Transforms above 2^20 points (rows of 2^14 points, columns of the rest).
# Exact integer convolutions for the checks: residues mod two NTT primes by
# the oracle's prime-generic NTT, recombined in uint64 (coefficients below
# 2^59, enough for residues of a prime below 2^16 over up to 2^26 terms).
_Q1, _G1, _Q2, _G2 = 2013265921, 31, 469762049, 3


def _exact_conv_mod(a, b, p, n_out):
    """(a * b)[0:n_out] mod p for uint32 residues of a small prime p."""
    N = 1 << (len(a) + len(b) - 1).bit_length()
    res = []
    for q, g in ((_Q1, _G1), (_Q2, _G2)):
        fa = np.zeros(N, np.uint64)
        fb = np.zeros(N, np.uint64)
        fa[: len(a)] = a
        fb[: len(b)] = b
        c = O.ntt_generic((O.ntt_generic(fa, q, g) * O.ntt_generic(fb, q, g)) % np.uint64(q), q, g, invert=True)
        res.append(c[:n_out].astype(np.uint64))
    inv = pow(_Q1, -1, _Q2)
    t = ((res[1] + np.uint64(_Q2) - res[0] % np.uint64(_Q2)) % np.uint64(_Q2)) * np.uint64(inv) % np.uint64(_Q2)
    return ((res[0] + np.uint64(_Q1) * t) % np.uint64(p)).astype(np.uint32)


@pytest.mark.parametrize("p,la,lb", [(P, 1100000, 1050000), (3323, 1200000, 1000000)])
def test_poly_mat_mul_above_2_20_points(p, la, lb):
    """Products whose transforms have 2^22 points: mod the NTT prime itself
    (against the oracle's product) and mod 3323, the reference's default
    prime, through the three CRT primes (against an exact convolution)."""
    eng = _eng()
    print(f"pmm p={p} la={la} lb={lb}", flush=True)
    rng = np.random.default_rng(la + p)
    A = rng.integers(0, p, (1, 2, la)).astype(np.uint32)
    B = rng.integers(0, p, (2, 1, lb)).astype(np.uint32)
    lo, hi = 1000, la + lb - 1
    got = eng.poly_mat_mul(A, B, lo, hi, p)
    if p == P:
        want = O.poly_mat_mul(A, B, lo, hi, p)
    else:
        want = np.zeros((1, 1, hi - lo), np.uint32)
        acc = np.zeros(hi, np.uint64)
        for r in range(2):
            acc = (acc + _exact_conv_mod(A[0, r], B[r, 0], p, hi)) % np.uint64(p)
        want[0, 0] = acc[lo:].astype(np.uint32)
    assert np.array_equal(got, want), np.argwhere(got != want)[:5]


@pytest.mark.parametrize("p", [P, 3323])
def test_pm_basis_order_above_2_20(p):
    """b = 1, d = 1,100,000: the root products need 2^21-point transforms
    (and, mod 3323, the CRT primes at that size).  Checked against what fixes
    the basis: u^T M(x) G(x) = 0 mod x^d (an exact convolution), det M(z) =
    +-z^(-sum delta), and every row degree bounded by -delta."""
    eng = _eng()
    b, d = 1, 1100000
    print(f"pm_basis b={b} d={d} p={p}", flush=True)
    rng = np.random.default_rng(d + p)
    seq = rng.integers(0, p, (d + 2, 1)).astype(np.uint32)
    M, delta = eng.pm_basis(seq, b, d, p)
    D = M.shape[2]
    assert M.shape == (2, 2, d + 1)
    u = rng.integers(0, p, 2)
    a = _left_combine(u, M, p)  # [2][D]: u^T M
    g0 = seq[:D, 0].astype(np.uint32)  # G = [S(x); 1]
    if p == P:
        N = 1 << (2 * D - 1).bit_length()
        fa = np.zeros(N, np.uint32)
        fb = np.zeros(N, np.uint32)
        fa[:D] = a[0]
        fb[:D] = g0
        c = O.ntt((O.ntt(fa).astype(np.uint64) * O.ntt(fb).astype(np.uint64) % np.uint64(P)).astype(np.uint32),
                  invert=True)[:d].astype(np.uint64)
    else:
        c = _exact_conv_mod(a[0].astype(np.uint32), g0, p, d).astype(np.uint64)
    c = (c + a[1][:d]) % np.uint64(p)  # + (u^T M)[1] * 1
    assert not c.any(), np.nonzero(c)[0][:5]
    z = int(rng.integers(2, p))
    Mz = np.zeros((2, 2), np.uint64)
    zp = np.ones(D, np.uint64)
    for k in range(1, D):
        zp[k] = zp[k - 1] * np.uint64(z) % np.uint64(p)
    for i in range(2):
        for j in range(2):
            Mz[i, j] = int((M[i, j].astype(np.uint64) * zp % np.uint64(p)).sum() % np.uint64(p))
    det = int((Mz[0, 0] * Mz[1, 1] + np.uint64(p) * np.uint64(p) - Mz[0, 1] * Mz[1, 0]) % np.uint64(p))
    e = int(-delta.sum())
    assert det in (pow(z, e, p), (-pow(z, e, p)) % p)
    nzmax = np.where(M.any(axis=1), np.arange(D)[None, :], -1).max(axis=1)
    assert np.all(nzmax <= -delta)


_EDGE = synth.edge_shapes(P)


@pytest.mark.parametrize("b", [1, 2, 4])
@pytest.mark.parametrize("name", list(_EDGE))
def test_rank_degenerate_shapes(name, b):
    """Degenerate inputs end to end: no nonzeros at all, one row, one column,
    1 x 1, b above n.  The sequence is the oracle's bit for bit and rr_rank
    (order L - 1) equals the oracle's pipeline, which here equals dense
    elimination."""
    eng = _eng()
    A = synth.csr_from_dense(_EDGE[name])
    dr, dc, V = synth.preconditioners(3, A.m, A.n, b, P)
    L = eng.seq_len(A.m, A.n, b)
    print(f"{name} b={b} L={L}", flush=True)
    want = O.krylov(A, P, b, dr, dc, V, 0, L)
    with eng.Krylov(A, P, b, dr, dc, V) as K:
        assert np.array_equal(K.run(L), want)
    r_oracle = O.rank_from_basis(*O.pm_basis(want, b, L - 1, P), b, P)
    got, _ = eng.rank(A, P, b, dr, dc, V, L, L - 1)
    assert got == r_oracle == O.dense_rank(A.to_dense(P), P)
