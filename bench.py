"""Benchmark: Krylov iterations per second of the block-Wiedemann sequence on
one MI355X, on the graph-complex-like workload (200000 x 180000, 12 distinct
Zipf(1.1) columns per row, entries +-1/+-2 mod p = 2013265921, b = 32).

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one Krylov iteration: one sparse pass (alternately by A and by A^T)
plus the b x b projection, producing one sequence matrix S_t.  The default K
is the workload's full sequence length (2 * 180000 / 32 + 8 = 11258 entries),
so the timed region is the whole Krylov phase of one rank computation, with
the matrix and vectors already resident in HBM.

Prints one JSON line.  `roofline` prices the dominant kernel (the fused
SpMM + projection pass) against HBM peak using algorithmic bytes per pass
(int32 columns, u32 values, each array touched once) and its average
duration from HIP events on the engine's stream; `cpu_baseline` times the
CPU oracle (a restatement of the reference's Rust path) on a bounded prefix
of the same workload on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "traffic_k_pass_graph_complex.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(A, p, b, dr, dc, V, iters=256):
    """Oracle (restated reference CPU path, OpenMP) on a prefix of the same
    workload: iterations/s from the difference of two runs (setup cancels)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    threads = max(1, min(threads, os.cpu_count() or 1, 16))
    t0 = time.perf_counter()
    O.krylov(A, p, b, dr, dc, V, 0, 1, nthreads=threads)
    t1 = time.perf_counter()
    O.krylov(A, p, b, dr, dc, V, 0, 1 + iters, nthreads=threads)
    t2 = time.perf_counter()
    per_iter = ((t2 - t1) - (t1 - t0)) / iters
    return {"value": round(1.0 / per_iter, 3), "unit": "iters/s", "cores": threads, "kind": "port",
            "sample": f"S_1..S_{iters} of the same 200000x180000 b=32 workload (oracle/krylov_oracle.c, "
                      f"{threads} OpenMP threads; setup cancelled by differencing two runs)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="timed iterations (default: full sequence)")
    ap.add_argument("--warmup", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--m", type=int, default=200000)
    ap.add_argument("--n", type=int, default=180000)
    ap.add_argument("--b", type=int, default=32)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the engine does not shard "
              f"a single matrix across GPUs", file=sys.stderr)
        sys.exit(2)

    from rust_rank_amd import engine, synth
    p, b = synth.P_DEFAULT, args.b
    t0 = time.perf_counter()
    A = synth.graph_complex(m=args.m, n=args.n)
    dr, dc, V = synth.preconditioners(2, A.m, A.n, b, p)
    L = synth.seq_len_plain(min(A.m, A.n), b)
    steps = args.steps or L
    log(f"workload {A.m}x{A.n} nnz={A.nnz} b={b} L={L}; generated in {time.perf_counter() - t0:.1f}s")

    K = engine.Krylov(A, p, b, dr, dc, V)
    K.run(max(args.warmup, 1))
    K.set_profiling(True)
    t1 = time.perf_counter()
    K.run(steps)
    t2 = time.perf_counter()
    st = K.stats()
    K.close()
    wall = t2 - t1
    iters_s = steps / wall
    passes = st["passes_a"] + st["passes_at"]
    pass_ms = (st["ms_a"] + st["ms_at"]) / max(passes, 1)
    bytes_pass = (st["bytes_a"] * st["passes_a"] + st["bytes_at"] * st["passes_at"]) / max(passes, 1)
    achieved = bytes_pass / (pass_ms * 1e-3) / 1e9
    b_iter = 8 * A.nnz + (4 * b + 2) * (A.m + A.n)
    log(f"timed {steps} iterations in {wall:.3f}s -> {iters_s:.1f} iters/s; k_pass avg {pass_ms * 1e3:.2f} us, "
        f"fixup {st['ms_fixup'] / max(passes, 1) * 1e3:.2f} us/pass, reduce {st['ms_reduce']:.2f} ms total, "
        f"stream total {st['ms_total']:.1f} ms")

    traffic = None
    if os.path.exists(PROFILE_TRAFFIC):
        with open(PROFILE_TRAFFIC) as f:
            traffic = json.load(f).get("bytes_per_launch")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": f"k_pass<{1 << (b - 1).bit_length()}, packed>",
                "bytes_per_launch": int(bytes_pass), "avg_launch_us": round(pass_ms * 1e3, 3)}
    cpu = None if args.no_cpu else cpu_baseline(A, p, b, dr, dc, V)
    out = {
        "metric": "Krylov iters/s + HBM GB/s; batched ranks/s & p50/p99 vs plain, same GPU",
        "value": round(iters_s, 2), "unit": "iters/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / steps, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32", "data": "synthetic (seeded SplitMix64)",
        "config": {"workload": f"graph_complex {A.m}x{A.n} 12 nnz/row Zipf(1.1) +-1/+-2 p=2013265921 b={b}",
                   "seq_len": L, "nnz": A.nnz, "hbm_gbs_algorithmic": round(b_iter * iters_s / 1e9, 1)},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
