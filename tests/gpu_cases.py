"""Inputs shared by several GPU test files, with their expected results from
the oracle (computed on the host when a test asks for them)."""
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
