"""Pins the CPU oracle (oracle/liboracle.so) before it is trusted as the checker:
against the reference's own compiled C++ (golden vectors in tests/golden,
written by tests/golden/make_golden.py), against closed forms (dense brute
force, DFT, determinants), and against invariants that catch dropped terms,
swapped indices or transposed operands."""
import os

import numpy as np
import pytest

import pyoracle as O
from dense_ref import krylov_dense, matmul_mod
from rust_rank_amd import synth

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
P = 2013265921
P61 = 2305843009146585089


def gold(name):
    return np.load(os.path.join(G, name))


# ------------------------------------------------------------------ NTT
def test_ntt_matches_reference_p61():
    d = gold("ntt_p61.npz")
    for i in range(4):
        a = d[f"in{i}"]
        assert np.array_equal(O.ntt_generic(a, P61, 3), d[f"fwd{i}"])
        assert np.array_equal(O.ntt_generic(d[f"fwd{i}"], P61, 3, invert=True), d[f"inv{i}"])
        assert np.array_equal(d[f"inv{i}"], a)


@pytest.mark.parametrize("n", [1, 2, 8, 64, 512])
def test_ntt_is_the_dft_mod_p(n):
    rng = np.random.default_rng(n)
    a = rng.integers(0, P, n).astype(np.uint32)
    got = O.ntt(a)
    w = pow(31, (P - 1) // n, P)  # root 31 (src/ntt.rs:51-53)
    want = [sum(int(a[j]) * pow(w, i * j, P) for j in range(n)) % P for i in range(n)]
    assert got.tolist() == want
    assert np.array_equal(O.ntt(got, invert=True), a)


def test_ntt_convolution_theorem():
    rng = np.random.default_rng(3)
    n = 4096
    a = np.zeros(n, np.uint32)
    b = np.zeros(n, np.uint32)
    a[:1000] = rng.integers(0, P, 1000)
    b[:1500] = rng.integers(0, P, 1500)
    fa, fb = O.ntt(a), O.ntt(b)
    c = O.ntt(((fa.astype(np.uint64) * fb) % P).astype(np.uint32), invert=True)
    want = np.zeros(n, dtype=object)
    ai = [int(x) for x in a[:1000]]
    bi = [int(x) for x in b[:1500]]
    for i in (0, 1, 777, 1999, 2498):
        want[i] = sum(ai[j] * bi[i - j] for j in range(max(0, i - 1499), min(i, 999) + 1)) % P
        assert int(c[i]) == want[i]


# ------------------------------------------------------------------ products
def test_poly_mat_mul_matches_reference():
    d = gold("poly_mat_mul.npz")
    for i in range(4):
        p, b0, b1 = (int(x) for x in d[f"p{i}"])
        A, B, Cr = d[f"A{i}"], d[f"B{i}"], d[f"C{i}"]
        got = O.poly_mat_mul(A, B[:, :, b0:b1], 0, Cr.shape[2], p)
        assert np.array_equal(got, Cr)


def test_poly_mat_mul_ntt_equals_schoolbook_big_p():
    rng = np.random.default_rng(5)
    A = rng.integers(0, P, (3, 4, 90)).astype(np.uint32)
    B = rng.integers(0, P, (4, 2, 70)).astype(np.uint32)
    a = O.poly_mat_mul(A, B, 10, 150, P, use_ntt=True)
    b = O.poly_mat_mul(A, B, 10, 150, P, use_ntt=False)
    assert np.array_equal(a, b)
    # one coefficient by hand
    i, j, k = 2, 1, 57
    want = sum(int(A[i, r, s]) * int(B[r, j, k - s]) for r in range(4) for s in range(max(0, k - 69), min(k, 89) + 1)) % P
    assert int(a[i, j, k - 10]) == want


# ------------------------------------------------------------------ dense linear algebra
def test_lsp_inverse_rank_match_reference():
    d = gold("linalg.npz")
    i = 0
    while f"a{i}" in d:
        p = int(d[f"p{i}"][0])
        a = d[f"a{i}"]
        L, S, perm, _ = O.lsp(a, p)
        assert np.array_equal(L, d[f"L{i}"])
        assert np.array_equal(S, d[f"S{i}"])
        # the reference's permutation matrix has P[i, perm[i]] = 1 (modular_linalg.h:345-348)
        Pm = np.zeros_like(L)
        Pm[np.arange(len(perm)), perm] = 1
        assert np.array_equal(Pm, d[f"P{i}"])
        inv = O.matrix_inverse(a, p)
        if d[f"inv{i}"].size:
            assert np.array_equal(inv, d[f"inv{i}"])
        else:
            assert inv is None
        assert O.modular_rank(a, p) == int(d[f"rank{i}"][0])
        # invariant: L S P^T... = A, with S's columns in pivot order
        LS = matmul_mod(L, S, p)
        A_re = np.zeros_like(LS)
        A_re[:, perm] = LS
        assert np.array_equal(A_re.astype(np.uint32), a % p)
        i += 1


# ------------------------------------------------------------------ PM-basis
def test_m_basis_matches_reference():
    d = gold("m_basis.npz")
    i = 0
    while f"G{i}" in d:
        p = int(d[f"p{i}"][0])
        M, dout = O.m_basis(d[f"G{i}"], d[f"din{i}"], p)
        assert np.array_equal(M, d[f"M{i}"]), i
        assert np.array_equal(dout, d[f"dout{i}"]), i
        i += 1


def test_pm_basis_and_rank_match_reference():
    d = gold("pm_basis.npz")
    i = 0
    while f"seq{i}" in d:
        p, b, dd, L = (int(x) for x in d[f"meta{i}"])
        basis, delta = O.pm_basis(d[f"seq{i}"], b, dd, p)
        assert np.array_equal(basis, d[f"basis{i}"]), i
        assert np.array_equal(delta, d[f"delta{i}"]), i
        assert O.rank_from_basis(basis, delta, b, p) == int(d[f"rank{i}"][0]), i
        i += 1


def _eval_pm(M, z, p):
    """M(z) for a [r][c][len] polynomial matrix (Horner)."""
    acc = np.zeros(M.shape[:2], np.uint64)
    for k in range(M.shape[2] - 1, -1, -1):
        acc = (acc * np.uint64(z) + M[:, :, k]) % np.uint64(p)
    return acc


def _det_mod(a, p):
    a = [[int(x) % p for x in row] for row in a]
    n, det = len(a), 1
    for c in range(n):
        piv = next((r for r in range(c, n) if a[r][c]), None)
        if piv is None:
            return 0
        if piv != c:
            a[c], a[piv] = a[piv], a[c]
            det = -det
        det = det * a[c][c] % p
        inv = pow(a[c][c], p - 2, p)
        for r in range(c + 1, n):
            f = a[r][c] * inv % p
            if f:
                a[r] = [(x - f * y) % p for x, y in zip(a[r], a[c])]
    return det % p


@pytest.mark.parametrize("b,d", [(2, 37), (4, 64), (3, 51)])
def test_pm_basis_invariants_big_p(b, d):
    """At p = 2013265921 (no reference run is exact there): the basis M is an order-d
    approximant of G = [S; I] (rows of M G vanish mod x^d), det M(z) = +-z^(-sum delta),
    and every row degree is bounded by -delta."""
    rng = np.random.default_rng(b * 100 + d)
    Pn = b * (b + 1) // 2
    seq = rng.integers(0, P, (d + 3, Pn)).astype(np.uint32)
    M, delta = O.pm_basis(seq, b, d, P)
    Gm = np.zeros((2 * b, b, d + 1), np.uint32)
    iu = np.triu_indices(b)
    for k in range(d + 1):
        S = np.zeros((b, b), np.uint32)
        S[iu] = seq[k]
        S = S + np.triu(S, 1).T
        Gm[:b, :, k] = S
    Gm[b:, :, 0] = np.eye(b, dtype=np.uint32)
    R = O.poly_mat_mul(M, Gm, 0, d, P)
    assert not R.any()
    z = 123456789
    det = _det_mod(_eval_pm(M, z, P), P)
    e = int(-delta.sum())
    assert det in (pow(z, e, P), (-pow(z, e, P)) % P)
    for i in range(2 * b):
        nz = np.nonzero(M[i].any(axis=0))[0]
        assert nz.size == 0 or nz.max() <= -delta[i]


def test_pm_basis_split_independent():
    """Non-power-of-two d through the generalised split equals the reference's power-of-two
    recursion truncated... checked as: basis(d) for d = 2^k equals the reference's, and
    basis(d1 + d2) = basis_(residual)(d2) * basis(d1) by construction; here we check the
    d=2^k result against a purely sequential chain of order-1 leaves."""
    rng = np.random.default_rng(9)
    b, d, p = 2, 16, 97
    seq = rng.integers(0, p, (d + 1, 3)).astype(np.uint32)
    M, delta = O.pm_basis(seq, b, d, p)
    # sequential: M_k on the constant term of the current residual
    Gm = np.zeros((2 * b, b, d + 1), np.uint32)
    iu = np.triu_indices(b)
    for k in range(d + 1):
        S = np.zeros((b, b), np.uint32)
        S[iu] = seq[k]
        Gm[:b, :, k] = S + np.triu(S, 1).T
    Gm[b:, :, 0] = np.eye(b, dtype=np.uint32)
    acc = np.zeros((2 * b, 2 * b, 1), np.uint32)
    acc[:, :, 0] = np.eye(2 * b, dtype=np.uint32)
    dl = np.zeros(2 * b, np.int64)
    for k in range(d):
        R = O.poly_mat_mul(acc, Gm, k, k + 1, p)[:, :, 0]
        Mk, dl = O.m_basis(R, dl, p)
        acc = O.poly_mat_mul(Mk, acc, 0, acc.shape[2] + 1, p)
    assert np.array_equal(acc[:, :, : d + 1], M)
    assert np.array_equal(dl, delta)


# ------------------------------------------------------------------ Krylov
@pytest.mark.parametrize("p,b,m,n", [(P, 1, 30, 25), (P, 4, 120, 100), (P, 3, 90, 110), (3323, 2, 60, 40)])
def test_krylov_closed_form(p, b, m, n):
    A = synth.uniform_sparse(b + m, 1, m, n, 1, 6, p, small=False)
    dr, dc, V = synth.preconditioners(b + n, m, n, b, p)
    L = 2 * min(m, n) // b + 8
    seq, x_last = O.krylov(A, p, b, dr, dc, V, 0, L, want_x=True)
    want, xs = krylov_dense(A.to_dense(p), dr, dc, V, p, L)
    assert np.array_equal(seq, want)
    assert np.array_equal(x_last, xs[(L - 1) // 2].astype(np.uint32))


def test_krylov_resume_and_hankel():
    p, b, m, n = P, 4, 200, 150
    A = synth.uniform_sparse(4, 1, m, n, 3, 7, p, small=True)
    dr, dc, V = synth.preconditioners(4, m, n, b, p)
    full, x20 = O.krylov(A, p, b, dr, dc, V, 0, 41, want_x=True)
    # resume from x_10 at t = 20 and at t = 21 (src/main.rs:425-443, the .wdm resume)
    _, x10 = O.krylov(A, p, b, dr, dc, V, 0, 21, want_x=True)
    tail = O.krylov(A, p, b, dr, dc, x10, 20, 41)
    assert np.array_equal(tail, full[20:])
    tail2 = O.krylov(A, p, b, dr, dc, x10, 21, 41)
    assert np.array_equal(tail2, full[21:])
    # Hankel: x_a . x_b depends only on a + b, so S_{2k+1} from (x_{k-1}, x_{k+2}) too
    _, xs = krylov_dense(A.to_dense(p), dr, dc, V, p, 30)
    iu = np.triu_indices(b)
    k = 6
    G = matmul_mod(xs[k - 1].T.copy(), xs[k + 2], p)
    assert np.array_equal(((G + G.T) * np.uint64((p + 1) // 2) % np.uint64(p))[iu], full[2 * k + 1])


def test_spmm_transpose_scale_on_reference_fixture():
    d = gold("matrix_txt.npz")
    m, n = (int(x) for x in d["shape"])
    rp = d["row_ptr"].astype(np.uint64)
    col = d["col"].astype(np.uint32)
    val = d["val"]
    rpt, colt, valt = O.transpose(m, n, rp, col, val)
    assert np.array_equal(rpt, d["row_ptr_t"].astype(np.uint64))
    assert np.array_equal(colt, d["col_t"].astype(np.uint32))
    assert np.array_equal(valt, d["val_t"])
    A = synth.Csr(m, n, rp, col, val)
    dense = A.to_dense(P)
    assert O.dense_rank(dense, P) == int(d["rank"][0])
    # D_r A D_c and D_c A^T as the reference scales them
    va = val.copy()
    O.lib().orc_scale_rows(m, rp, va, d["dr"], P)
    O.lib().orc_scale_cols(len(va), col, va, d["dc"], P)
    assert np.array_equal(va, d["a_scaled"])
    vt = valt.copy()
    O.lib().orc_scale_rows(n, rpt, vt, d["dc"], P)
    assert np.array_equal(vt, d["at_scaled"])


def test_rank_end_to_end_matches_dense_ge():
    """Krylov -> PM-basis (full order) -> rank formula equals dense Gaussian elimination."""
    for seed, (m, n, dep, b) in enumerate([(80, 70, 5, 2), (150, 120, 9, 4), (64, 80, 3, 4)]):
        A = synth.uniform_sparse(seed, 5, m, n, 2, 6, P, small=True)
        A = synth.plant_dependent_columns(synth.Rng(seed, 6), A, dep, P, small=True)
        dr, dc, V = synth.preconditioners(seed, m, n, b, P)
        L = 2 * min(m, n) // b + 8
        seq = O.krylov(A, P, b, dr, dc, V, 0, L)
        basis, delta = O.pm_basis(seq, b, L - 1, P)
        assert O.rank_from_basis(basis, delta, b, P) == O.dense_rank(A.to_dense(P), P)


def test_plain_small_rank():
    """The 2000 x 1800 plain workload: rank by BW equals dense GE (b = 8, L = 458)."""
    A = synth.plain_small()
    b = 8
    dr, dc, V = synth.preconditioners(1, A.m, A.n, b, P)
    L = synth.seq_len_plain(min(A.m, A.n), b)
    assert L == 458
    seq = O.krylov(A, P, b, dr, dc, V, 0, L)
    basis, delta = O.pm_basis(seq, b, L - 1, P)
    ge = O.dense_rank(A.to_dense(P), P)
    assert O.rank_from_basis(basis, delta, b, P) == ge


def test_batch_matrices_parallel_equals_serial():
    """The batched workload generated by worker processes is the same
    matrices as generated one by one (the bench and the GPU tests rely on it)."""
    par = synth.batch_matrices(4, 20, workers=2)
    ser = [synth.batch_matrix(4, i) for i in range(20)]
    for A, B in zip(par, ser):
        assert (A.m, A.n) == (B.m, B.n)
        assert np.array_equal(A.row_ptr, B.row_ptr)
        assert np.array_equal(A.col, B.col)
        assert np.array_equal(A.val, B.val)
