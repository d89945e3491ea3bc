"""Writes tests/golden/graph_complex_digest.json: the SHA-256 of the whole
Krylov sequence of the 200000 x 180000 graph-complex workload (b = 32,
11258 entries S_0..S_11257, t-major [L][528] little-endian u32) computed by
the CPU oracle (oracle/krylov_oracle.c, a restatement of src/main.rs:88-216),
and the rank the oracle's PM-basis of order 11257 and rank formula
(oracle/pmbasis_oracle.cpp: src/sigma_basis.rs:9-152, src/blockbmtest.rs:574-600)
give on it.  The GPU test compares its own sequence and rr_rank with these.

    make -C oracle && python tests/golden/make_graph_complex_digest.py

About 5 minutes of Krylov and 10-20 minutes of PM-basis on 8 cores.
"""
import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402
import pyoracle as O  # noqa: E402
from rust_rank_amd import synth  # noqa: E402

P = synth.P_DEFAULT


def main():
    b = 32
    A = synth.graph_complex()
    dr, dc, V = synth.preconditioners(2, A.m, A.n, b, P)
    L = synth.seq_len_plain(min(A.m, A.n), b)
    t0 = time.time()
    seq = O.krylov(A, P, b, dr, dc, V, 0, L)
    t1 = time.time()
    print(f"oracle Krylov: {L} entries in {t1 - t0:.0f}s", flush=True)
    digest = hashlib.sha256(np.ascontiguousarray(seq, dtype="<u4").tobytes()).hexdigest()
    prefix = {str(t): hashlib.sha256(np.ascontiguousarray(seq[:t], dtype="<u4").tobytes()).hexdigest()
              for t in (48, 1024, 5629)}
    M, delta = O.pm_basis(seq, b, L - 1, P)
    rank = O.rank_from_basis(M, delta, b, P)
    t2 = time.time()
    print(f"oracle PM-basis of order {L - 1} and rank {rank} in {t2 - t1:.0f}s", flush=True)
    out = {"workload": "graph_complex 200000x180000 12 nnz/row Zipf(1.1) +-1/+-2 p=2013265921 b=32 "
                       "(synth.graph_complex(), synth.preconditioners(2, ...))",
           "seq_len": L, "pairs": b * (b + 1) // 2, "sha256_seq": digest, "sha256_prefix": prefix,
           "order": L - 1, "rank": int(rank), "delta_sum": int(delta.sum()),
           "s_last_first_words": [int(x) for x in seq[-1][:8]]}
    with open(os.path.join(HERE, "graph_complex_digest.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
