"""GPU: the batched mode (rr_batch_rank) gives, for every matrix, the rank the
single-matrix path gives, and dense Gaussian elimination confirms it."""
import numpy as np
import pytest

import pyoracle as O
from rust_rank_amd import synth
from gpu_cases import p3323_cases, split_row_cases

P = 2013265921
pytestmark = pytest.mark.gpu


def _eng():
    from rust_rank_amd import engine
    assert engine.device_count() >= 1, "no GPU visible"
    return engine


def test_batch_degenerate_shapes_beside_ordinary_matrices():
    """No nonzeros, 1 x 1, one row, one column and b above n, batched with
    ordinary matrices: every rank equals the single path and dense
    elimination."""
    eng = _eng()
    mats = [synth.csr_from_dense(D) for D in synth.edge_shapes(P).values()]
    mats += [synth.batch_matrix(23, i, rows_lo=200, rows_hi=600) for i in range(3)]
    bs = [(1, 2, 4)[i % 3] for i in range(len(mats))]
    pre = [synth.preconditioners(700 + i, A.m, A.n, b, P) for i, (A, b) in enumerate(zip(mats, bs))]
    print(f"batch of {len(mats)}: shapes {[(A.m, A.n) for A in mats]}, b {bs}", flush=True)
    ranks, _ = eng.batch_rank(mats, P, bs, [x[0] for x in pre], [x[1] for x in pre], [x[2] for x in pre])
    for i, A in enumerate(mats):
        single, _ = eng.rank(A, P, bs[i], *pre[i])
        assert ranks[i] == single == O.dense_rank(A.to_dense(P), P), (i, ranks[i], single)


def test_batch_other_prime_matches_oracle():
    eng = _eng()
    p, cases = p3323_cases()
    print(f"p={p}: shapes {[(c[0].m, c[0].n, c[1]) for c in cases]}", flush=True)
    ranks, _ = eng.batch_rank([c[0] for c in cases], p, [c[1] for c in cases], [c[2] for c in cases],
                              [c[3] for c in cases], [c[4] for c in cases], [c[5] for c in cases])
    for i, (A, b, dr, dc, V, L, want) in enumerate(cases):
        assert ranks[i] == want == O.dense_rank(A.to_dense(p), p), (i, ranks[i], want)


def test_batch_lane_width_64():
    """b = 64 and b = 48 (lane width 64) in a batch: the batched passes fuse the
    projection into the 64-lane pass, where the single path projects in a
    kernel of its own; the ranks agree with each other and with dense
    elimination."""
    eng = _eng()
    mats = [synth.batch_matrix(71, i, rows_lo=300, rows_hi=900) for i in range(5)]
    bs = [64, 48, 64, 16, 64]
    pre = [synth.preconditioners(950 + i, A.m, A.n, b, P) for i, (A, b) in enumerate(zip(mats, bs))]
    print(f"batch of {len(mats)}: shapes {[(A.m, A.n) for A in mats]}, b {bs}", flush=True)
    ranks, _ = eng.batch_rank(mats, P, bs, [x[0] for x in pre], [x[1] for x in pre], [x[2] for x in pre])
    for i, A in enumerate(mats):
        single, _ = eng.rank(A, P, bs[i], *pre[i])
        assert ranks[i] == single == O.dense_rank(A.to_dense(P), P), (i, bs[i], ranks[i], single)


def test_batch_general_values():
    """Matrices with values uniform in [1, p) (not small signed residues: the
    passes take the general-value form) at several lane widths, batched: the
    ranks agree with the single path and dense elimination."""
    eng = _eng()
    shapes = [(400, 360), (500, 520), (300, 300), (650, 600), (420, 380)]
    bs = [4, 16, 32, 64, 3]
    mats = [synth.uniform_sparse(75, i, m, n, 3, 9, P, small=False) for i, (m, n) in enumerate(shapes)]
    mats[0] = synth.plant_dependent_columns(synth.Rng(75, 99), mats[0], 12, P, False)
    pre = [synth.preconditioners(990 + i, A.m, A.n, b, P) for i, (A, b) in enumerate(zip(mats, bs))]
    ranks, _ = eng.batch_rank(mats, P, bs, [x[0] for x in pre], [x[1] for x in pre], [x[2] for x in pre])
    for i, A in enumerate(mats):
        single, _ = eng.rank(A, P, bs[i], *pre[i])
        assert ranks[i] == single == O.dense_rank(A.to_dense(P), P), (i, bs[i], ranks[i], single)


def test_batch_split_rows_match_oracle():
    """Matrices whose A^T rows are cut into segments (long Zipf columns), batched
    beside an ordinary matrix: each job's segment sums meet in its own
    accumulator rows within shared launches; ranks equal the oracle's
    pipeline and the single path."""
    eng = _eng()
    cases = list(split_row_cases())
    A3 = synth.batch_matrix(79, 0, rows_lo=800, rows_hi=1200)
    cases.append((A3, 16) + tuple(synth.preconditioners(1130, A3.m, A3.n, 16, P)) +
                 (synth.seq_len_plain(min(A3.m, A3.n), 16), O.dense_rank(A3.to_dense(P), P)))
    ranks, _ = eng.batch_rank([c[0] for c in cases], P, [c[1] for c in cases], [c[2] for c in cases],
                              [c[3] for c in cases], [c[4] for c in cases], [c[5] for c in cases])
    for i, (A, b, dr, dc, V, L, want) in enumerate(cases):
        single, _ = eng.rank(A, P, b, dr, dc, V, L, L - 1)
        assert ranks[i] == single == want, (i, ranks[i], single, want)


def test_batch_matches_single_and_dense_ge():
    eng = _eng()
    mats, bs, drs, dcs, Vs = [], [], [], [], []
    for i in range(12):
        A = synth.batch_matrix(21, i, rows_lo=150, rows_hi=900)
        b = (4, 8, 16)[i % 3]
        dr, dc, V = synth.preconditioners(100 + i, A.m, A.n, b, P)
        mats.append(A); bs.append(b); drs.append(dr); dcs.append(dc); Vs.append(V)
    print(f"batch of {len(mats)}: rows {[A.m for A in mats]}", flush=True)
    ranks, stats = eng.batch_rank(mats, P, bs, drs, dcs, Vs)
    print(stats, flush=True)
    for i, A in enumerate(mats):
        single, _ = eng.rank(A, P, bs[i], drs[i], dcs[i], Vs[i])
        ge = O.dense_rank(A.to_dense(P), P)
        assert ranks[i] == single == ge, (i, ranks[i], single, ge)


def test_batch_sequences_bit_identical_to_oracle():
    """Same matrices alone and in a mix: the ranks come from the same S_t; here the
    S_t of one batched matrix are compared with the oracle through its rank at a
    short order (different orders exercise different sequence prefixes)."""
    eng = _eng()
    A = synth.batch_matrix(5, 1, rows_lo=200, rows_hi=300)
    b = 4
    dr, dc, V = synth.preconditioners(7, A.m, A.n, b, P)
    L = 2 * min(A.m, A.n) // b + 8
    seq = O.krylov(A, P, b, dr, dc, V, 0, L)
    for d in (L - 1, L // 2, 17):
        basis, delta = O.pm_basis(seq, b, d, P)
        want = O.rank_from_basis(basis, delta, b, P)
        other = synth.batch_matrix(5, 2, rows_lo=100, rows_hi=400)
        dr2, dc2, V2 = synth.preconditioners(8, other.m, other.n, 8, P)
        ranks, _ = eng.batch_rank([other, A], P, [8, b], [dr2, dr], [dc2, dc], [V2, V], [0, L], [0, d])
        assert ranks[1] == want, (d, ranks[1], want)


def test_batch_streamed_pm_basis_matches_oracle():
    """Orders of at least 1024 start their PM-basis while the sequence is still
    being produced (the leftmost chain of the recursion reads only leading
    coefficients); the rank at each such order equals the oracle's."""
    eng = _eng()
    A = synth.batch_matrix(6, 3, rows_lo=2600, rows_hi=2700)
    b = 4
    dr, dc, V = synth.preconditioners(9, A.m, A.n, b, P)
    L = 2 * min(A.m, A.n) // b + 8
    assert L - 1 >= 1024
    seq = O.krylov(A, P, b, dr, dc, V, 0, L)
    other = synth.batch_matrix(6, 4, rows_lo=100, rows_hi=400)
    dr2, dc2, V2 = synth.preconditioners(10, other.m, other.n, 8, P)
    for d in (L - 1, 1024 + 37):
        basis, delta = O.pm_basis(seq, b, d, P)
        want = O.rank_from_basis(basis, delta, b, P)
        ranks, _ = eng.batch_rank([other, A], P, [8, b], [dr2, dr], [dc2, dc], [V2, V], [0, L], [0, d])
        print(f"m={A.m} n={A.n} L={L} d={d}: rank {ranks[1]} oracle {want}", flush=True)
        assert ranks[1] == want, (d, ranks[1], want)


def test_batch_many_matrices_match_single():
    """48 matrices: set-up on worker threads overlaps the device work, so
    requests join the tick loop at different times; every rank still equals
    the single-matrix path's."""
    eng = _eng()
    mats, bs, pre = [], [], []
    for i in range(48):
        A = synth.batch_matrix(33, i, rows_lo=100, rows_hi=3000)
        b = (8, 16, 4, 32)[i % 4]
        mats.append(A)
        bs.append(b)
        pre.append(synth.preconditioners(300 + i, A.m, A.n, b, P))
    print(f"batch of {len(mats)}: rows {min(A.m for A in mats)}..{max(A.m for A in mats)}", flush=True)
    ranks, stats = eng.batch_rank(mats, P, bs, [x[0] for x in pre], [x[1] for x in pre], [x[2] for x in pre])
    print(stats, flush=True)
    for i, A in enumerate(mats):
        single, _ = eng.rank(A, P, bs[i], *pre[i])
        assert ranks[i] == single, (i, ranks[i], single)


def test_batch_rejects_invalid_matrix_and_recovers():
    """A value out of range in one matrix of a batch: the call fails with an
    error naming that matrix (its set-up threads joined), and the next batch
    on the same device runs normally."""
    eng = _eng()
    mats, pre = [], []
    for i in range(20):
        A = synth.batch_matrix(44, i, rows_lo=100, rows_hi=800)
        mats.append(A)
        pre.append(synth.preconditioners(400 + i, A.m, A.n, 8, P))
    bad = mats[7]
    val = bad.val.copy()
    val[3] = P  # not a residue
    mats[7] = synth.Csr(bad.m, bad.n, bad.row_ptr, bad.col, val)
    with pytest.raises(RuntimeError, match="matrix 7"):
        eng.batch_rank(mats, P, [8] * 20, [x[0] for x in pre], [x[1] for x in pre], [x[2] for x in pre])
    mats[7] = bad
    # a start-vector word >= p, found on the device during that matrix's set-up
    V5 = pre[5][2].copy()
    V5[3, 1] = P
    with pytest.raises(RuntimeError, match=r"matrix 5: V\[%d\] >= p" % (3 * 8 + 1)):
        eng.batch_rank(mats, P, [8] * 20, [x[0] for x in pre], [x[1] for x in pre],
                       [x[2] if i != 5 else V5 for i, x in enumerate(pre)])
    ranks, _ = eng.batch_rank(mats[:6], P, [8] * 6, [x[0] for x in pre[:6]], [x[1] for x in pre[:6]],
                              [x[2] for x in pre[:6]])
    for i in range(6):
        single, _ = eng.rank(mats[i], P, 8, *pre[i])
        assert ranks[i] == single, (i, ranks[i], single)


def test_cached_device_memory_reuse_and_release():
    """Buffers freed by one call are reused by the next (the cache waits for
    the device before handing them out) and can be returned to the system;
    results stay equal across reuse and release."""
    eng = _eng()
    mats, pre = [], []
    for i in range(10):
        A = synth.batch_matrix(55, i, rows_lo=200, rows_hi=2000)
        mats.append(A)
        pre.append(synth.preconditioners(500 + i, A.m, A.n, 16, P))
    args = (P, [16] * 10, [x[0] for x in pre], [x[1] for x in pre], [x[2] for x in pre])
    first, _ = eng.batch_rank(mats, *args)
    again, _ = eng.batch_rank(mats, *args)  # served from the cache
    eng.release_device_memory()
    fresh, _ = eng.batch_rank(mats, *args)
    assert list(first) == list(again) == list(fresh)
    singles = [eng.rank(A, P, 16, *pre[i])[0] for i, A in enumerate(mats)]
    assert list(first) == singles


def test_batch_beyond_one_staging_slot():
    """2000 tiny matrices in one rr_batch_rank: a chunk's tick descriptors
    (16 ticks x 2000 requests) exceed one staging slot and go down in several
    staged launches; every rank equals dense Gaussian elimination, and a
    sample equals the single-matrix path."""
    eng = _eng()
    n = 2000
    mats, bs, drs, dcs, Vs = [], [], [], [], []
    for i in range(n):
        A = synth.batch_matrix(41, i, rows_lo=40, rows_hi=120)
        b = (4, 8)[i % 2]
        dr, dc, V = synth.preconditioners(7000 + i, A.m, A.n, b, P)
        mats.append(A); bs.append(b); drs.append(dr); dcs.append(dc); Vs.append(V)
    print(f"batch of {n} matrices, rows {min(A.m for A in mats)}..{max(A.m for A in mats)}", flush=True)
    ranks, stats = eng.batch_rank(mats, P, bs, drs, dcs, Vs)
    print(stats, flush=True)
    for i, A in enumerate(mats):
        assert ranks[i] == O.dense_rank(A.to_dense(P), P), i
    for i in range(0, n, 250):
        assert ranks[i] == eng.rank(mats[i], P, bs[i], drs[i], dcs[i], Vs[i])[0], i


def test_autotuned_widths_and_block_sizes_give_dense_ranks():
    """rr_autotune picks b and the column block size su per matrix from its
    probes; a batch run with those choices (rr_batch_rank_tuned) gives every
    matrix the rank dense elimination gives.  Rows of 8 nonzeros pad least in
    blocks of 4, rows of 12 in blocks of 3: the mix gives the model both
    cases, and forcing the other block size gives the same ranks."""
    eng = _eng()
    mats = []
    for i in range(9):
        k = (8, 12, 10)[i % 3]
        A = synth.uniform_sparse(40 + i, 7, 300 + 37 * i, 270 + 33 * i, k, k, P, small=True)
        mats.append(synth.plant_dependent_columns(synth.Rng(40 + i, 9), A, 5 + i, P, small=True))
    bs, sus, pred = eng.autotune(mats, P)
    print(f"autotune: b {bs.tolist()} su {sus.tolist()} pred ms {np.round(pred, 3).tolist()}", flush=True)
    assert set(sus.tolist()) <= {3, 4} and set(bs.tolist()) <= {8, 16, 32}
    drs, dcs, Vs = [], [], []
    for i, A in enumerate(mats):
        dr, dc, V = synth.preconditioners(200 + i, A.m, A.n, int(bs[i]), P)
        drs.append(dr); dcs.append(dc); Vs.append(V)
    ranks, _ = eng.batch_rank(mats, P, bs, drs, dcs, Vs, sus=sus)
    other, _ = eng.batch_rank(mats, P, bs, drs, dcs, Vs, sus=7 - sus)
    for i, A in enumerate(mats):
        ge = O.dense_rank(A.to_dense(P), P)
        assert ranks[i] == other[i] == ge, (i, ranks[i], other[i], ge)
