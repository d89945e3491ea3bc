"""Inputs shared by several GPU test files, with their expected results from
the oracle (computed on the host when a test asks for them)."""
import functools

import pyoracle as O
from rust_rank_amd import synth


def p3323_cases():
    """Six matrices mod p = 3323 (the reference's default prime), b in {2, 4, 8}:
    with many jobs in a round, most PM-basis products go through the three
    CRT primes in shared launches.  Each with its rank from the oracle's
    pipeline (Krylov loop, PM-basis of order L - 1, rank formula)."""
    p = 3323
    out = []
    for i in range(6):
        A = synth.batch_matrix(61, i, p=p, rows_lo=300, rows_hi=1500)
        b = (4, 8, 2)[i % 3]
        dr, dc, V = synth.preconditioners(900 + i, A.m, A.n, b, p)
        L = synth.seq_len_plain(min(A.m, A.n), b)
        want = O.rank_from_basis(*O.pm_basis(O.krylov(A, p, b, dr, dc, V, 0, L), b, L - 1, p), b, p)
        out.append((A, b, dr, dc, V, L, want))
    return p, out


@functools.lru_cache(maxsize=1)
def split_row_cases():
    """Two graph-complex-like matrices (12 nonzeros per row, columns drawn
    Zipf(1.1)): the most frequent columns hold thousands of entries, so A^T
    has rows long enough to be cut into segments that several wavefronts add
    into one accumulator row.  Each with its rank from the oracle's pipeline."""
    P = synth.P_DEFAULT
    out = []
    for seed, m, n, b in ((11, 6000, 5000, 16), (12, 4000, 3500, 8)):
        A = synth.graph_complex(seed=seed, m=m, n=n, k=12)
        dr, dc, V = synth.preconditioners(1100 + seed, A.m, A.n, b, P)
        L = synth.seq_len_plain(min(A.m, A.n), b)
        seq = O.krylov(A, P, b, dr, dc, V, 0, L)
        out.append((A, b, dr, dc, V, L, O.rank_from_basis(*O.pm_basis(seq, b, L - 1, P), b, P)))
    return tuple(out)
