"""GPU parity of the Krylov engine (rr_krylov_* through the C ABI) against the
CPU oracle on identical seeded inputs: every S_t must be bit-identical."""
import sys

import numpy as np
import pytest

import pyoracle as O
from rust_rank_amd import synth

P = 2013265921
pytestmark = pytest.mark.gpu


def _engine():
    from rust_rank_amd import engine
    assert engine.device_count() >= 1, "no GPU visible"
    return engine


def _case(A, p, b, seed, L, t_have=0):
    eng = _engine()
    dr, dc, V = synth.preconditioners(seed, A.m, A.n, b, p)
    print(f"case m={A.m} n={A.n} nnz={A.nnz} b={b} p={p} L={L}", flush=True)
    want, x_want = O.krylov(A, p, b, dr, dc, V, 0, L, want_x=True)
    with eng.Krylov(A, p, b, dr, dc, V) as K:
        got = K.run(L)
        curv, t = K.state()
        on_device = K.stats()["setup_on_device"]
    assert t == L
    assert np.array_equal(got, want), np.argwhere(got != want)[:5]
    assert np.array_equal(curv, x_want)
    return on_device


@pytest.mark.parametrize("b", [1, 2, 3, 4, 8, 16, 32, 48, 64])
def test_general_values(b):
    A = synth.uniform_sparse(b, 1, 300, 260, 1, 9, P, small=False)
    _case(A, P, b, b, 2 * 260 // b + 8)


@pytest.mark.parametrize("b", [4, 8, 32, 64])
def test_small_values_packed(b):
    A = synth.batch_matrix(3, b, rows_lo=400, rows_hi=600)
    _case(A, P, b, b + 1, 40)


def test_small_prime():
    A = synth.uniform_sparse(9, 1, 200, 150, 2, 6, 3323, small=False)
    _case(A, 3323, 8, 9, 45)


def test_long_rows_split():
    """A^T rows far above the split threshold (Zipf head): segments finished
    in the pass; such a matrix is laid out on the host."""
    A = synth.graph_complex(seed=5, m=40000, n=2000, k=12)
    assert int(np.bincount(A.col).max()) > 4096
    assert _case(A, P, 8, 5, 12) == 0


@pytest.mark.parametrize("b,small", [(4, True), (16, False), (64, False)])
def test_device_setup_edge_cases(b, small):
    """Set-up on the device (entry checks, transpose, column-block streams):
    empty rows and columns, repeated columns within a row, rows of every
    length up to 40, packed and general values; the sequence and the state
    equal the oracle's."""
    rng = np.random.default_rng(b)
    m, n = 700, 500
    lens = rng.integers(0, 41, m)
    lens[rng.choice(m, 60, replace=False)] = 0  # empty rows
    live = rng.choice(n, n - 50, replace=False)  # 50 empty columns
    cols = [np.sort(rng.choice(live, k)) for k in lens]  # with repeats
    col = np.concatenate(cols).astype(np.uint32)
    if small:
        val = rng.choice(np.array([1, 2, 7, P - 1, P - 2, P - 8], np.uint32), len(col))
    else:
        val = rng.integers(1, P, len(col), dtype=np.uint64).astype(np.uint32)
    row_ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    A = synth.Csr(m, n, row_ptr, col, val)
    assert _case(A, P, b, 40 + b, 30) == 1


def test_long_rows_split_b64():
    """Lane width 64 (projection in its own kernel) with split A^T rows finished in-pass."""
    A = synth.graph_complex(seed=5, m=40000, n=2000, k=12)
    _case(A, P, 64, 5, 12)


@pytest.mark.timeout(900)
def test_large_single_full_size_prefix():
    """The 5000000 x 4500000, 20 nonzeros per row, b = 64 workload itself
    (general values, 5000 planted dependent columns): S_0..S_5, three passes
    by A and two by A^T, and the state x_2, against the oracle."""
    A = synth.large_single()
    _case(A, P, 64, 2, 6)


@pytest.mark.timeout(900)
def test_block_vectors_above_2_gib():
    """8,600,000 x 8,500,000 with 2 nonzeros per row at b = 33 (stride 64):
    each block vector is 2.2 GB, past 2^31 bytes, so the gathers' byte offsets
    use the whole 32-bit range.  S_0..S_2 and the state against the oracle."""
    A = synth.uniform_sparse(11, 1, 8_600_000, 8_500_000, 2, 2, P, small=False)
    assert A.m * 64 * 4 > 1 << 31
    _case(A, P, 33, 11, 3)


def test_graph_complex_full_size_prefix():
    """The benchmark workload itself (200000 x 180000, b = 32): 48 entries,
    the A^T passes with hundreds of split-row segments finished in-pass."""
    A = synth.graph_complex()
    _case(A, P, 32, 2, 48)


def test_split_rows_many_passes():
    """Every A^T row split (40000 x 2000, ~240 nonzeros per column), 96
    entries at b = 16: the segment hand-off through atomics, repeatedly."""
    A = synth.graph_complex(seed=6, m=40000, n=2000, k=12)
    _case(A, P, 16, 6, 96)


def test_plain_small_full_sequence():
    A = synth.plain_small()
    L = synth.seq_len_plain(min(A.m, A.n), 8)
    _case(A, P, 8, 1, L)


def test_resume_matches_continuous():
    eng = _engine()
    A = synth.uniform_sparse(2, 1, 500, 400, 3, 8, P, small=False)
    b = 4
    dr, dc, V = synth.preconditioners(2, A.m, A.n, b, P)
    with eng.Krylov(A, P, b, dr, dc, V) as K:
        first = K.run(21)
        curv, t = K.state()
        rest = K.run(20)
    assert t == 21
    with eng.Krylov(A, P, b, dr, dc, curv, t_have=21) as K2:
        rest2 = K2.run(20)
    assert np.array_equal(rest, rest2)
    want = O.krylov(A, P, b, dr, dc, V, 0, 41)
    assert np.array_equal(np.vstack([first, rest]), want)


def test_bad_inputs_raise():
    eng = _engine()
    A = synth.uniform_sparse(2, 1, 50, 40, 3, 8, P, small=False)
    dr, dc, V = synth.preconditioners(2, A.m, A.n, 4, P)
    with pytest.raises(eng.RankEngineError):
        eng.Krylov(A, P, 65, dr, dc, np.zeros((A.n, 65), np.uint32))
    with pytest.raises(eng.RankEngineError):
        eng.Krylov(A, 2013265920, 4, dr, dc, V)  # not prime
    bad = A.val.copy()
    bad[7] = P
    B = synth.Csr(A.m, A.n, A.row_ptr, A.col, bad)
    with pytest.raises(eng.RankEngineError, match=r"val\[7\]"):
        eng.Krylov(B, P, 4, dr, dc, V)
    badc = A.col.copy()
    badc[-1] = A.n
    C = synth.Csr(A.m, A.n, A.row_ptr, badc, A.val)
    with pytest.raises(eng.RankEngineError, match=r"col\[%d\]" % (A.nnz - 1)):
        eng.Krylov(C, P, 4, dr, dc, V)
    rp = A.row_ptr.copy()
    rp[5], rp[6] = rp[6], rp[5] - 1
    D = synth.Csr(A.m, A.n, rp, A.col, A.val)
    with pytest.raises(eng.RankEngineError):
        eng.Krylov(D, P, 4, dr, dc, V)
    # start vectors are checked on the device: the first word >= p is named,
    # at b = B (uploaded as they are) and at b < B (padded there)
    for bb in (4, 3):
        Vb = np.ascontiguousarray(V[:, :bb])
        Vb[11, 2] = P + 5
        Vb[30, 0] = P
        with pytest.raises(eng.RankEngineError, match=r"curv\[%d\] >= p" % (11 * bb + 2)):
            eng.Krylov(A, P, bb, dr, dc, Vb)
    with pytest.raises(eng.RankEngineError):
        eng.Krylov(A, P, 4, dr, dc, V, t_have=2)


def test_resume_from_wdm_checkpoint(tmp_path):
    """A .wdm checkpoint written mid-sequence (src/main.rs:186-202) and read
    back resumes to the same entries as the continuous run, with the entry
    count the reference's loop would append (src/main.rs:105-131)."""
    eng = _engine()
    A = synth.uniform_sparse(4, 1, 600, 500, 3, 9, P, small=False)
    b, max_nlen = 8, 41
    dr, dc, V = synth.preconditioners(4, A.m, A.n, b, P)
    with eng.Krylov(A, P, b, dr, dc, V) as K:
        first = K.run(21)
        curv, t = K.state()
        rest = K.run(20)
    path = str(tmp_path / "ck.wdm")
    eng.wdm_save(path, A.m, A.n, P, dr, dc, V, curv, first)
    w = eng.wdm_load(path)
    assert w["len"] == t == 21 and np.array_equal(w["V"], V) and np.array_equal(w["seq"], first)
    todo = eng.seq_resume_count(w["len"], max_nlen)
    assert todo == 20
    with eng.Krylov(A, P, b, w["row_precond"], w["col_precond"], w["curv"], t_have=w["len"]) as K2:
        again = K2.run(todo)
    assert np.array_equal(again, rest)
    assert np.array_equal(np.vstack([first, again]), O.krylov(A, P, b, dr, dc, V, 0, 41))


@pytest.mark.timeout(600)
def test_graph_complex_full_sequence_digest_and_rank():
    """The whole 11258-entry sequence of the 200000 x 180000, b = 32 workload
    (every split-row hand-off of every A^T pass included) against the
    oracle's: SHA-256 of the t-major array and of its prefixes; the device
    PM-basis of order 11257 against the oracle's degree shifts and rank; and
    rr_rank against that rank (tests/golden/graph_complex_digest.json, written
    by tests/golden/make_graph_complex_digest.py)."""
    import hashlib
    import json
    import os
    eng = _engine()
    g = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                    "graph_complex_digest.json")))
    A = synth.graph_complex()
    b = 32
    dr, dc, V = synth.preconditioners(2, A.m, A.n, b, P)
    L = synth.seq_len_plain(min(A.m, A.n), b)
    assert L == g["seq_len"]
    with eng.Krylov(A, P, b, dr, dc, V) as K:
        seq = K.run(L)
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a, dtype="<u4").tobytes()).hexdigest()
    for t, h in g["sha256_prefix"].items():
        assert sha(seq[: int(t)]) == h, t
    assert sha(seq) == g["sha256_seq"]
    basis, delta = eng.pm_basis(seq, b, L - 1, P)
    assert int(delta.sum()) == g["delta_sum"]
    assert eng.rank_from_basis(basis, delta, b, P) == g["rank"]
    del basis
    assert eng.rank(A, P, b, dr, dc, V, L, L - 1)[0] == g["rank"]
