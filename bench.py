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
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "traffic_k_spmm_graph_complex.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(A, p, b, dr, dc, V, iters=1024):
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


def batched_and_served(eng, synth, p, nmat, rate_frac, seed=4, plain=True):
    """Batched form: nmat matrices of the batched workload (rows log-uniform in
    [1000, 50000], cols 0.9 rows, 6-16 nnz/row, +-1/+-2, 1-5 % dependent
    columns), b per matrix from the autotuner, solved in one rr_batch_rank, vs
    the same matrices solved one at a time by the plain GPU path (rr_rank).
    Served form: the same matrices as a Poisson stream at rate_frac x the
    measured batched capacity through rr_server; FIFO plain latencies are
    computed from the measured plain service times on the same arrival times."""
    import threading
    t0 = time.perf_counter()
    mats = synth.batch_matrices(seed, nmat)
    bs, pred, probe = eng.autotune_b(mats, p)
    pre = [synth.preconditioners(5000 + i, A.m, A.n, int(b), p) for i, (A, b) in enumerate(zip(mats, bs))]
    log(f"batched: {nmat} matrices, rows {min(A.m for A in mats)}..{max(A.m for A in mats)}, "
        f"b histogram {dict(zip(*np.unique(bs, return_counts=True)))}, setup {time.perf_counter() - t0:.1f}s")
    if not plain:  # the batched call alone (for kernel traces)
        t2 = time.perf_counter()
        ranks, st = eng.batch_rank(mats, p, bs, [x[0] for x in pre], [x[1] for x in pre], [x[2] for x in pre])
        t_batch = time.perf_counter() - t2
        log(f"batched {t_batch:.2f}s (plain solves skipped)")
        return {"matrices": nmat, "ranks_per_s": round(nmat / t_batch, 3), "wall_s": round(t_batch, 3),
                "ms_krylov": round(st["ms_krylov"], 1), "ms_pm": round(st["ms_pm"], 1)}, None
    # batched: one call (first, while the device is as fresh as for the Krylov leg)
    t2 = time.perf_counter()
    ranks, st = eng.batch_rank(mats, p, bs, [x[0] for x in pre], [x[1] for x in pre], [x[2] for x in pre])
    t_batch = time.perf_counter() - t2
    # plain: one matrix at a time
    plain_ms, plain_rank = [], []
    t1 = time.perf_counter()
    for (A, b, (dr, dc, V)) in zip(mats, bs, pre):
        r, ms = eng.rank(A, p, int(b), dr, dc, V)
        plain_rank.append(r)
        plain_ms.append(ms)
    t_plain = time.perf_counter() - t1
    assert list(ranks) == plain_rank, "batched ranks differ from plain ranks"
    batched = {"matrices": nmat, "ranks_per_s": round(nmat / t_batch, 3), "plain_ranks_per_s": round(nmat / t_plain, 3),
               "speedup_vs_plain": round(t_plain / t_batch, 2), "wall_s": round(t_batch, 3),
               "plain_wall_s": round(t_plain, 3), "ticks": st["ticks"], "pm_rounds": st["pm_rounds"],
               "ms_krylov": round(st["ms_krylov"], 1), "ms_pm": round(st["ms_pm"], 1),
               "b_hist": {int(k): int(v) for k, v in zip(*np.unique(bs, return_counts=True))},
               "autotune_probe": probe, "ranks_equal_plain": True}
    log(f"batched {t_batch:.2f}s vs plain {t_plain:.2f}s -> {t_plain / t_batch:.1f}x")
    # served: Poisson arrivals at rate_frac x batched capacity
    lam = rate_frac * nmat / t_batch
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / lam, nmat))
    results = {}
    srv = eng.Server(p, 200 << 30)
    stop = threading.Event()

    def poller():
        while not stop.is_set() or len(results) < nmat:
            for r in srv.poll(wait_ms=50):
                results[r["id"]] = r
            if len(results) >= nmat:
                break

    th = threading.Thread(target=poller)
    th.start()
    t3 = time.perf_counter()
    ids = []
    for i, (A, b, (dr, dc, V)) in enumerate(zip(mats, bs, pre)):
        delay = t3 + arrivals[i] - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        ids.append(srv.submit(A, int(b), dr, dc, V))
    stop.set()
    th.join(timeout=600)
    t_served = time.perf_counter() - t3
    sst = srv.stats()
    srv.close()
    lat = np.array([results[i]["latency_ms"] for i in ids])
    served_ok = all(results[i]["rank"] == plain_rank[k] for k, i in enumerate(ids))
    # FIFO plain on the same arrivals: one solve at a time in arrival order
    fin, fifo = 0.0, []
    for a, ms in zip(arrivals * 1e3, plain_ms):
        fin = max(fin, a) + ms
        fifo.append(fin - a)
    fifo = np.array(fifo)
    served = {"requests": nmat, "offered_rate_per_s": round(lam, 3), "ranks_per_s": round(nmat / t_served, 3),
              "p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
              "fifo_plain_p50_ms": round(float(np.percentile(fifo, 50)), 2),
              "fifo_plain_p99_ms": round(float(np.percentile(fifo, 99)), 2),
              "fifo_plain_ranks_per_s": round(nmat / (fin / 1e3), 3),
              "max_active": sst["max_active"], "hbm_budget_gb": 200, "ranks_equal_plain": bool(served_ok)}
    log(f"served: p50 {served['p50_ms']} ms p99 {served['p99_ms']} ms vs FIFO plain p50 {served['fifo_plain_p50_ms']} "
        f"p99 {served['fifo_plain_p99_ms']}")
    return batched, served


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="timed iterations (default: full sequence)")
    ap.add_argument("--warmup", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-pm", action="store_true", help="skip the PM-basis and rank of the sequence")
    ap.add_argument("--m", type=int, default=200000)
    ap.add_argument("--n", type=int, default=180000)
    ap.add_argument("--b", type=int, default=0, help="block width (default: 32, or 64 for --workload large)")
    ap.add_argument("--workload", choices=["graph_complex", "large"], default="graph_complex",
                    help="large: the 5000000 x 4500000, 20 nnz/row single matrix (Krylov only, bounded steps)")
    ap.add_argument("--batched", type=int, default=512, help="matrices in the batched/served comparison (0: skip)")
    ap.add_argument("--rate-frac", type=float, default=0.7, help="served offered load / batched capacity")
    ap.add_argument("--no-plain", action="store_true",
                    help="batched call only: skip the plain solves and the served form (for kernel traces)")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MB",
                    help="only run the plain copy kernel on MB megabytes (counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calib:
        from rust_rank_amd import engine
        ms, gbs = engine.copy_probe(args.copy_calib << 20, 20)
        print(json.dumps({"copy_bytes": args.copy_calib << 20, "avg_ms": ms, "read_write_gbs": gbs}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the engine does not shard "
              f"a single matrix across GPUs", file=sys.stderr)
        sys.exit(2)

    from rust_rank_amd import engine, synth
    large = args.workload == "large"
    with_pm = not (large or args.no_pm)
    p, b = synth.P_DEFAULT, args.b or (64 if large else 32)
    t0 = time.perf_counter()
    if large:  # Krylov rate on a bounded stretch: the whole sequence is about 140k iterations
        A = synth.large_single()
        desc = f"large_single {A.m}x{A.n} 20 nnz/row uniform, values uniform in [1,p), deficit 5000, p=2013265921 b={b}"
    else:
        A = synth.graph_complex(m=args.m, n=args.n)
        desc = f"graph_complex {A.m}x{A.n} 12 nnz/row Zipf(1.1) +-1/+-2 p=2013265921 b={b}"
    dr, dc, V = synth.preconditioners(2, A.m, A.n, b, p)
    L = synth.seq_len_plain(min(A.m, A.n), b)
    steps = args.steps or (256 if large else L)
    log(f"workload {A.m}x{A.n} nnz={A.nnz} b={b} L={L}; generated in {time.perf_counter() - t0:.1f}s")

    K = engine.Krylov(A, p, b, dr, dc, V)
    seq = [K.run(max(args.warmup, 1))]
    t1 = time.perf_counter()
    seq.append(K.run(steps))  # timed: no per-pass events
    t2 = time.perf_counter()
    have = sum(len(x) for x in seq)
    if have < L and with_pm:  # the rest of the sequence for the PM-basis, untimed
        seq.append(K.run(L - have))
    seq = np.concatenate(seq)[:L]
    # kernel breakdown on a further stretch of the same sequence, with HIP
    # events around every launch on the engine's stream
    K.set_profiling(True)
    K.run(min(steps, 2000))
    st = K.stats()
    K.close()
    wall = t2 - t1
    iters_s = steps / wall
    passes = st["passes_a"] + st["passes_at"]
    pass_ms = (st["ms_a"] + st["ms_at"]) / max(passes, 1)
    bytes_pass = (st["bytes_a"] * st["passes_a"] + st["bytes_at"] * st["passes_at"]) / max(passes, 1)
    achieved = bytes_pass / (pass_ms * 1e-3) / 1e9
    b_iter = 8 * A.nnz + (4 * b + 2) * (A.m + A.n)
    log(f"pass by A avg {st['ms_a'] / max(st['passes_a'], 1) * 1e3:.2f} us, by A^T avg "
        f"{st['ms_at'] / max(st['passes_at'], 1) * 1e3:.2f} us")
    log(f"timed {steps} iterations in {wall:.3f}s -> {iters_s:.1f} iters/s; profiled stretch: k_spmm avg "
        f"{pass_ms * 1e3:.2f} us, "
        f"reduce {st['ms_reduce']:.2f} ms, stream total {st['ms_total']:.1f} ms")

    traffic = None
    if os.path.exists(PROFILE_TRAFFIC) and not large:
        with open(PROFILE_TRAFFIC) as f:
            traffic = json.load(f).get("bytes_per_launch")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": f"k_spmm<{1 << (b - 1).bit_length()}, {'general' if large else 'packed'}>",
                "bytes_per_launch": int(bytes_pass), "avg_launch_us": round(pass_ms * 1e3, 3)}
    # the rest of the rank pipeline on this sequence: device PM-basis of order
    # L - 1, then the rank formula
    pm = None
    if with_pm:
        t3 = time.perf_counter()
        basis, delta = engine.pm_basis(seq, b, L - 1, p)
        t4 = time.perf_counter()
        rank = engine.rank_from_basis(basis, delta, b, p)
        del basis
        pm = {"b": b, "order": L - 1, "ms": round((t4 - t3) * 1e3, 1), "rank": int(rank),
              "rank_pipeline_s": round(L / iters_s + (t4 - t3), 3)}
        log(f"PM-basis of order {L - 1}: {pm['ms']} ms; rank {rank}")
    cpu = None if (args.no_cpu or large) else cpu_baseline(A, p, b, dr, dc, V)
    batched = served = None
    if args.batched and not large:
        batched, served = batched_and_served(engine, synth, p, args.batched, args.rate_frac, plain=not args.no_plain)
    if pm is not None:
        # the same rank through rr_rank: set-up, sequence and PM-basis in one
        # call, the PM-basis running beside the passes on this workload (run
        # last: the batched figures above were measured without it before them)
        rr_rank, rr_ms = engine.rank(A, p, b, dr, dc, V, L, L - 1)
        assert int(rr_rank) == pm["rank"], (rr_rank, pm["rank"])
        pm["rr_rank_s"] = round(rr_ms / 1e3, 3)
    out = {
        "metric": "Krylov iters/s + HBM GB/s; batched ranks/s & p50/p99 vs plain, same GPU",
        "value": round(iters_s, 2), "unit": "iters/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / steps, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32", "data": "synthetic (seeded SplitMix64)",
        "config": {"workload": desc,
                   "seq_len": L, "nnz": A.nnz, "hbm_gbs_algorithmic": round(b_iter * iters_s / 1e9, 1)},
        "roofline": roofline, "cpu_baseline": cpu, "pm_basis": pm, "batched": batched, "served": served,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
