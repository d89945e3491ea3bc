"""Seeded synthetic inputs for the rank engine.

The reference draws its preconditioners and block V from an unseeded thread
RNG (src/matrices.rs:484-497, src/main.rs:464-473), so two runs never see the
same inputs.  Here every draw comes from a counter-based SplitMix64 stream
keyed by (seed, stream id), so the HIP engine and the CPU oracle can be fed
identical inputs and compared bit for bit.

Workloads (sizes as named):
  plain_small      2000 x 1800, 8 nnz per base row, 200 dependent rows, b = 8
  graph_complex    200000 x 180000, 12 distinct Zipf(1.1) columns per row, +-1/+-2, b = 32
  large_single     5,000,000 x 4,500,000, 20 uniform columns per row, 5000 dependent columns, b = 64
  batched          rows log-uniform in [1000, 50000], cols = 0.9 rows, 6-16 nnz per row,
                   +-1/+-2, 1-5 % planted dependent columns
  served           the same with rows log-uniform in [10^3, 10^6] (rehearsal: [10^3, 10^5])
                   and 8-20 nnz per row

Matrices are CSR with uint64 row pointers, uint32 column indices and uint32
values in [0, p) (negative entries stored as p - |v|).
"""
from __future__ import annotations

import dataclasses
import os

import numpy as np

P_DEFAULT = 2013265921  # 15 * 2^27 + 1

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


_PAR_MIN = 1 << 22  # work items below which the helpers below stay on one thread


def _blocks(n: int, align: int = 1):
    """[lo, hi) ranges splitting n items (multiples of align) over up to 16 threads."""
    t = max(1, min(16, os.cpu_count() or 1))
    step = max(align, (n + t - 1) // t // align * align or align)
    return [(lo, min(n, lo + step)) for lo in range(0, n, step)]


def _pmap(fn, ranges):
    """fn(lo, hi) over the ranges on a thread pool (numpy releases the GIL)."""
    if len(ranges) <= 1:
        return [fn(lo, hi) for lo, hi in ranges]
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=len(ranges)) as ex:
        return list(ex.map(lambda r: fn(*r), ranges))


class Rng:
    """SplitMix64 stream: output i = mix(key + (i + 1) * golden)."""

    def __init__(self, seed: int, stream: int):
        with np.errstate(over="ignore"):
            s = _mix(np.array([seed & 0xFFFFFFFFFFFFFFFF], dtype=np.uint64) + _GOLD)
            k = _mix(s ^ _mix(np.array([stream & 0xFFFFFFFFFFFFFFFF], dtype=np.uint64) + _GOLD))
        self.key = k[0]
        self.ctr = 0

    def u64(self, n: int) -> np.ndarray:
        out = np.empty(n, dtype=np.uint64)
        base = self.ctr + 1

        def fill(lo, hi):
            with np.errstate(over="ignore"):
                idx = np.arange(base + lo, base + hi, dtype=np.uint64)
                out[lo:hi] = _mix(self.key + idx * _GOLD)

        _pmap(fill, _blocks(n) if n >= _PAR_MIN else [(0, n)])
        self.ctr += n
        return out

    def integers(self, lo: int, hi: int, n: int) -> np.ndarray:
        """n integers in [lo, hi) (hi - lo < 2^32), multiply-shift mapping."""
        span = np.uint64(hi - lo)
        x = self.u64(n) >> np.uint64(32)
        return (x * span >> np.uint64(32)).astype(np.int64) + lo

    def uniform(self, n: int) -> np.ndarray:
        """n doubles in [0, 1)."""
        return (self.u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


@dataclasses.dataclass
class Csr:
    m: int
    n: int
    row_ptr: np.ndarray  # uint64, m + 1
    col: np.ndarray  # uint32, nnz
    val: np.ndarray  # uint32, nnz, in [0, p)

    @property
    def nnz(self) -> int:
        return int(self.row_ptr[-1])

    def to_dense(self, p: int) -> np.ndarray:
        d = np.zeros((self.m, self.n), dtype=np.uint64)
        rows = np.repeat(np.arange(self.m), np.diff(self.row_ptr.astype(np.int64)))
        np.add.at(d, (rows, self.col.astype(np.int64)), self.val.astype(np.uint64))
        return (d % np.uint64(p)).astype(np.uint32)


def _distinct_rows(cands: np.ndarray, k: np.ndarray | int):
    """For each row of cands (R x K), the first k distinct entries in draw order.
    Returns (flat selected values in row order, per-row count found)."""
    R, K = cands.shape
    if R * K >= _PAR_MIN:  # rows are independent: blocks of rows on threads
        kk = np.broadcast_to(np.asarray(k), (R,)) if np.ndim(k) else None
        parts = _pmap(lambda lo, hi: _distinct_rows_1(cands[lo:hi], k if kk is None else kk[lo:hi]), _blocks(R))
        return np.concatenate([v for v, _ in parts]), np.concatenate([f for _, f in parts])
    return _distinct_rows_1(cands, k)


def _distinct_rows_1(cands: np.ndarray, k: np.ndarray | int):
    R, K = cands.shape
    order = np.argsort(cands, axis=1, kind="stable")
    srt = np.take_along_axis(cands, order, axis=1)
    dup_sorted = np.zeros_like(srt, dtype=bool)
    dup_sorted[:, 1:] = srt[:, 1:] == srt[:, :-1]
    dup = np.empty_like(dup_sorted)
    np.put_along_axis(dup, order, dup_sorted, axis=1)
    keep = ~dup
    rank = np.cumsum(keep, axis=1)
    kk = np.broadcast_to(np.asarray(k).reshape(-1, 1) if np.ndim(k) else np.full((R, 1), k), (R, 1))
    sel = keep & (rank <= kk)
    found = sel.sum(axis=1)
    return cands[sel], found


def _sample_rows(rng: Rng, R: int, k, draw) -> list:
    """Per row, k (scalar or array) distinct columns from the sampler `draw(count)`."""
    kk = np.full(R, k, dtype=np.int64) if np.ndim(k) == 0 else np.asarray(k, dtype=np.int64)
    out_cols = [None] * R
    todo = np.arange(R)
    extra = 4
    while todo.size:
        K = int(kk[todo].max()) + extra
        c = draw(todo.size * K).reshape(todo.size, K)
        vals, found = _distinct_rows(c, kk[todo])
        ok = found == kk[todo]
        # split flat values back per row
        starts = np.concatenate([[0], np.cumsum(found)])
        for i in np.nonzero(ok)[0]:
            out_cols[todo[i]] = vals[starts[i]:starts[i + 1]]
        todo = todo[~ok]
        extra *= 2
    return out_cols


def _rows_to_csr(m: int, n: int, cols_per_row: list, vals_per_row: list) -> Csr:
    lens = np.array([len(c) for c in cols_per_row], dtype=np.uint64)
    rp = np.zeros(m + 1, dtype=np.uint64)
    np.cumsum(lens, out=rp[1:])
    col = np.concatenate(cols_per_row).astype(np.uint32) if m else np.zeros(0, np.uint32)
    val = np.concatenate(vals_per_row).astype(np.uint32) if m else np.zeros(0, np.uint32)
    return Csr(m, n, rp, col, val)


def _fixed_k_csr(m: int, n: int, k: np.ndarray, cols_flat: np.ndarray, vals_flat: np.ndarray) -> Csr:
    rp = np.zeros(m + 1, dtype=np.uint64)
    np.cumsum(k.astype(np.uint64), out=rp[1:])
    # sort columns inside each row (distinct within a row, so the order is unique)
    if m and np.all(k == k[0]):
        kk = int(k[0])
        c2, v2 = cols_flat.reshape(m, kk), vals_flat.reshape(m, kk)
        col = np.empty((m, kk), dtype=np.uint32)
        val = np.empty((m, kk), dtype=np.uint32)

        def srt(lo, hi):
            o2 = np.argsort(c2[lo:hi], axis=1, kind="stable")
            col[lo:hi] = np.take_along_axis(c2[lo:hi], o2, axis=1)
            val[lo:hi] = np.take_along_axis(v2[lo:hi], o2, axis=1)

        _pmap(srt, _blocks(m) if m * kk >= _PAR_MIN else [(0, m)])
        return Csr(m, n, rp, col.reshape(-1), val.reshape(-1))
    rows = np.repeat(np.arange(m, dtype=np.int64), k.astype(np.int64))
    order = np.argsort(rows * n + cols_flat.astype(np.int64), kind="stable")  # by row, then column
    return Csr(m, n, rp, cols_flat[order].astype(np.uint32), vals_flat[order].astype(np.uint32))


def _sample_fixed_k(rng: Rng, m: int, k: np.ndarray, draw) -> np.ndarray:
    """Flat array of per-row distinct columns (row-major), vectorised."""
    K = int(k.max()) + 8
    c = draw(m * K).reshape(m, K)
    vals, found = _distinct_rows(c, k)
    bad = np.nonzero(found != k)[0]
    if bad.size == 0:
        return vals
    # rare: redo the short rows one by one
    starts = np.concatenate([[0], np.cumsum(found)])
    pieces = []
    for r in range(m):
        if found[r] == k[r]:
            pieces.append(vals[starts[r]:starts[r + 1]])
        else:
            pieces.append(_sample_rows(rng, 1, int(k[r]), draw)[0])
    return np.concatenate(pieces)


def csr_from_dense(D: np.ndarray) -> Csr:
    """The nonzero entries of a dense matrix (values already in [0, p)) as CSR,
    row by row in column order."""
    rp, cols, vals = [0], [], []
    for i in range(D.shape[0]):
        nz = np.nonzero(D[i])[0]
        cols += nz.tolist()
        vals += D[i, nz].tolist()
        rp.append(len(cols))
    return Csr(int(D.shape[0]), int(D.shape[1]), np.array(rp, np.uint64), np.array(cols, np.uint32),
               np.array(vals, np.uint32))


def edge_shapes(p: int = P_DEFAULT, seed: int = 5) -> dict:
    """Degenerate dense matrices for end-to-end checks: no nonzeros, 1 x 1,
    one row, one column, a full-rank diagonal, 3 x 2."""
    rng = np.random.default_rng(seed)
    return {
        "zero 50x40": np.zeros((50, 40), np.uint64),
        "1x1": np.array([[5]], np.uint64),
        "row 1x30": rng.integers(0, p, (1, 30)).astype(np.uint64),
        "column 30x1": rng.integers(0, p, (30, 1)).astype(np.uint64),
        "diagonal 40": np.diag(rng.integers(1, p, 40)).astype(np.uint64),
        "3x2": rng.integers(0, p, (3, 2)).astype(np.uint64),
    }


def small_values(rng: Rng, nnz: int, p: int) -> np.ndarray:
    """+-1 (70 %) and +-2 (30 %), equal signs, stored mod p."""
    mag = np.where(rng.uniform(nnz) < 0.7, 1, 2)
    neg = rng.u64(nnz) & np.uint64(1)
    v = np.where(neg == 1, p - mag, mag)
    return v.astype(np.uint32)


def plant_dependent_columns(rng: Rng, A: Csr, k: int, p: int, small: bool) -> Csr:
    """Overwrite k columns with combinations of 2 (small) or 3 other columns, so
    rank(A) <= n - k.  Small matrices use coefficients +-1 (entries stay in [-4, 4])."""
    if k <= 0:
        return A
    n, m = A.n, A.m
    perm = np.argsort(rng.u64(n))
    dep = perm[:k]
    free = perm[k:]
    terms = 2 if small else 3
    src = free[rng.integers(0, free.size, k * terms)].reshape(k, terms)
    if small:
        coef = np.where(rng.u64(k * terms) & np.uint64(1), p - 1, 1).astype(np.uint64).reshape(k, terms)
    else:
        coef = rng.integers(1, p, k * terms).astype(np.uint64).reshape(k, terms)
    # Only the entries of the source columns move: collect them (each with
    # every (t, s) pair that reads its column), merge the new entries of each
    # (row, dependent column) mod p, and splice them into the kept rows.
    rows = np.repeat(np.arange(m, dtype=np.int64), np.diff(A.row_ptr.astype(np.int64)))
    col = A.col.astype(np.int64)
    is_dep = np.zeros(n, dtype=bool)
    is_dep[dep] = True
    cflat = src.reshape(-1)
    is_src = np.zeros(n, dtype=bool)
    is_src[cflat] = True
    sel = np.nonzero(is_src[col] & ~is_dep[col])[0]  # entries of the source columns (row order)
    srow, scol, sval = rows[sel], col[sel], A.val[sel].astype(np.uint64)
    so = np.argsort(scol, kind="stable")  # by column, each column in row order
    srow, scol, sval = srow[so], scol[so], sval[so]
    starts = np.searchsorted(scol, np.arange(n + 1))
    cnt = starts[cflat + 1] - starts[cflat]
    first = np.repeat(starts[cflat] - np.concatenate([[0], np.cumsum(cnt)[:-1]]), cnt)
    idx = first + np.arange(int(cnt.sum()))
    pair = np.repeat(np.arange(cflat.size), cnt)
    nr = srow[idx]
    nc = dep[pair // terms].astype(np.int64)
    nv = sval[idx] * coef.reshape(-1)[pair] % np.uint64(p)
    key = nr * n + nc
    o = np.argsort(key, kind="stable")
    key, nv = key[o], nv[o]
    uk, fst = np.unique(key, return_index=True)
    sums = np.add.reduceat(nv, fst) % np.uint64(p) if nv.size else nv
    nz = sums != 0
    uk, sums = uk[nz], sums[nz]
    keep = ~is_dep[col]
    krow, kcol, kval = rows[keep], col[keep], A.val[keep]
    kkey = krow * n + kcol  # ascending: rows in order, columns sorted within rows
    assert kkey.size < 2 or bool(np.all(kkey[1:] > kkey[:-1])), "rows must hold sorted, distinct columns"
    pos = np.searchsorted(kkey, uk) + np.arange(uk.size)  # final places of the new entries
    tot = kkey.size + uk.size
    isnew = np.zeros(tot, dtype=bool)
    isnew[pos] = True
    out_c = np.empty(tot, dtype=np.uint32)
    out_v = np.empty(tot, dtype=np.uint32)
    out_c[pos] = (uk % n).astype(np.uint32)
    out_v[pos] = sums.astype(np.uint32)
    out_c[~isnew] = kcol.astype(np.uint32)
    out_v[~isnew] = kval
    cnt_row = np.bincount(krow, minlength=m) + np.bincount((uk // n).astype(np.int64), minlength=m)
    rp = np.zeros(m + 1, dtype=np.uint64)
    np.cumsum(cnt_row, out=rp[1:])
    return Csr(m, n, rp, out_c, out_v)


def plain_small(seed: int = 1, p: int = P_DEFAULT, m_base: int = 1800, n: int = 1800,
                m_dep: int = 200, k: int = 8) -> Csr:
    """m_base base rows with k distinct uniform columns and values in [1, p-1],
    then m_dep rows that are combinations c1 r_a + c2 r_b + c3 r_c of 3 distinct
    base rows (c uniform in [1, p-1]), merged, exact zeros dropped."""
    rng = Rng(seed, 101)
    kk = np.full(m_base, k, dtype=np.int64)
    cols = _sample_fixed_k(rng, m_base, kk, lambda c: rng.integers(0, n, c))
    vals = rng.integers(1, p, m_base * k)
    base = _fixed_k_csr(m_base, n, kk, cols, vals)
    rows_c, rows_v = [], []
    rp = base.row_ptr.astype(np.int64)
    for _ in range(m_dep):
        src = _sample_rows(rng, 1, 3, lambda c: rng.integers(0, m_base, c))[0]
        coef = rng.integers(1, p, 3)
        acc = {}
        for r, cf in zip(src, coef):
            for e in range(rp[r], rp[r + 1]):
                c = int(base.col[e])
                acc[c] = (acc.get(c, 0) + int(cf) * int(base.val[e])) % p
        cs = sorted(c for c, v in acc.items() if v != 0)
        rows_c.append(np.array(cs, dtype=np.uint32))
        rows_v.append(np.array([acc[c] for c in cs], dtype=np.uint32))
    base_c = [base.col[rp[r]:rp[r + 1]] for r in range(m_base)]
    base_v = [base.val[rp[r]:rp[r + 1]] for r in range(m_base)]
    return _rows_to_csr(m_base + m_dep, n, base_c + rows_c, base_v + rows_v)


def zipf_sampler(rng: Rng, n: int, s: float):
    w = np.arange(1, n + 1, dtype=np.float64) ** (-s)
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    perm = np.argsort(rng.u64(n)).astype(np.int64)

    def draw(count):
        u = rng.uniform(count)
        r = np.minimum(np.searchsorted(cdf, u, side="right"), n - 1)
        return perm[r]

    return draw


def graph_complex(seed: int = 2, p: int = P_DEFAULT, m: int = 200000, n: int = 180000,
                  k: int = 12, s: float = 1.1) -> Csr:
    """m x n, k distinct columns per row from Zipf(s) over ranks mapped through a
    seeded column permutation, values +-1 (70 %) / +-2 (30 %)."""
    rng = Rng(seed, 202)
    draw = zipf_sampler(rng, n, s)
    kk = np.full(m, k, dtype=np.int64)
    cols = _sample_fixed_k(rng, m, kk, draw)
    vals = small_values(rng, m * k, p)
    return _fixed_k_csr(m, n, kk, cols, vals)


def uniform_sparse(seed: int, stream: int, m: int, n: int, k_lo: int, k_hi: int, p: int,
                   small: bool) -> Csr:
    """m x n with k uniform in [k_lo, k_hi] distinct uniform columns per row."""
    rng = Rng(seed, stream)
    kk = rng.integers(k_lo, k_hi + 1, m) if k_lo != k_hi else np.full(m, k_lo, dtype=np.int64)
    kk = np.minimum(kk, n)
    cols = _sample_fixed_k(rng, m, kk, lambda c: rng.integers(0, n, c))
    tot = int(kk.sum())
    vals = small_values(rng, tot, p) if small else rng.integers(1, p, tot).astype(np.uint32)
    return _fixed_k_csr(m, n, kk, cols, vals)


def large_single(seed: int = 3, p: int = P_DEFAULT, m: int = 5_000_000, n: int = 4_500_000,
                 k: int = 20, deficit: int = 5000) -> Csr:
    A = uniform_sparse(seed, 303, m, n, k, k, p, small=False)
    return plant_dependent_columns(Rng(seed, 304), A, deficit, p, small=False)


def batch_matrix(seed: int, index: int, p: int = P_DEFAULT, rows_lo: int = 1000, rows_hi: int = 50000,
                 k_lo: int = 6, k_hi: int = 16) -> Csr:
    """One matrix of the batched/served workloads: rows log-uniform in
    [rows_lo, rows_hi], cols = floor(0.9 rows), k_lo..k_hi nnz per row, +-1/+-2,
    1-5 % of rows planted as dependent columns."""
    rng = Rng(seed, 10_000 + index)
    u = rng.uniform(2)
    rows = int(np.exp(np.log(rows_lo) + u[0] * (np.log(rows_hi) - np.log(rows_lo))))
    cols = int(0.9 * rows)
    frac = 0.01 + 0.04 * u[1]
    A = uniform_sparse(seed, 20_000 + index, rows, cols, k_lo, k_hi, p, small=True)
    return plant_dependent_columns(Rng(seed, 30_000 + index), A, int(round(frac * rows)), p, small=True)


SERVED_ROWS_HI = 1_000_000     # the served workload's row range is [10^3, 10^6]
SERVED_REHEARSAL_ROWS_HI = 100_000  # the scaled rehearsal caps rows at 10^5


def served_matrix(seed: int, index: int, rows_hi: int = SERVED_ROWS_HI, p: int = P_DEFAULT) -> Csr:
    """One request of the served workload: rows log-uniform in [10^3, rows_hi],
    cols = floor(0.9 rows), 8-20 nnz per row, +-1/+-2, 1-5 % of rows planted
    as dependent columns (the batched generator with the served ranges)."""
    return batch_matrix(seed, index, p, rows_lo=1000, rows_hi=rows_hi, k_lo=8, k_hi=20)


def _batch_matrix_args(args):
    return batch_matrix(*args)


def _served_matrix_args(args):
    return served_matrix(*args)


def batch_matrices(seed: int, count: int, workers: int = 0) -> list:
    """batch_matrix(seed, i) for i < count, generated by up to `workers`
    processes (0: one per CPU, at most 16); the same matrices as one by one."""
    workers = workers or min(16, os.cpu_count() or 1)
    if workers <= 1 or count < 16:
        return [batch_matrix(seed, i) for i in range(count)]
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    with ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
        return list(ex.map(_batch_matrix_args, [(seed, i) for i in range(count)], chunksize=8))


def served_matrices(seed: int, count: int, rows_hi: int = SERVED_ROWS_HI, workers: int = 0) -> list:
    """served_matrix(seed, i, rows_hi) for i < count, on up to `workers` processes."""
    workers = workers or min(16, os.cpu_count() or 1)
    if workers <= 1 or count < 16:
        return [served_matrix(seed, i, rows_hi) for i in range(count)]
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    with ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
        return list(ex.map(_served_matrix_args, [(seed, i, rows_hi) for i in range(count)], chunksize=4))


def preconditioners(seed: int, m: int, n: int, b: int, p: int = P_DEFAULT):
    """D_r uniform in [1, p) (m), D_c uniform in [1, p) (n), V uniform in [0, p)
    (n x b, row-interleaved), as src/main.rs:464-469 draws them."""
    rng = Rng(seed, 909)
    dr = rng.integers(1, p, m).astype(np.uint32)
    dc = rng.integers(1, p, n).astype(np.uint32)
    V = rng.integers(0, p, n * b).astype(np.uint32).reshape(n, b)
    return dr, dc, V


def seq_len_plain(n_min: int, b: int) -> int:
    """Sequence length 2 min(m, n) / b + 8 used by the plain workloads."""
    return 2 * n_min // b + 8
