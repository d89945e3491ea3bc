"""Benchmark: Krylov iterations per second of the block-Wiedemann sequence on
one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

The headline is the largest single-GPU workload: a 5,000,000 x 4,500,000
matrix with 20 uniform columns per row and values uniform in [1, p)
(p = 2013265921, 5000 planted dependent columns, about 1.0e8 nonzeros) at
block width b = 64.  A step is one Krylov iteration: one sparse pass
(alternately by A and by A^T) plus the b x b projection, producing one
sequence matrix S_t.  The timed region is K iterations after W untimed ones,
with A, A^T and the block vectors resident in HBM and the device sequence
buffer already sized, ended by the copy of the K new entries to the host.

Prints one JSON line.  `roofline` prices the dominant kernel (the sparse
pass k_spmm) against HBM peak with the algorithmic bytes of one pass (int32
columns, u32 values, each array touched once) over its average duration from
events the dispatches stamp on the engine's stream; `cpu_baseline` times the
CPU oracle (a restatement of the reference's Rust path) on a prefix of the
same workload on this host.  Extra objects: the 200000 x 180000 graph-complex
workload (b = 32) over its whole sequence with its PM-basis and rank, the
2000 x 1800 plain workload end to end on the GPU and through the oracle on
the host, the 512-matrix batched form against plain solves, and the served
form.

    python bench.py --full-rank     # config-3 rank end to end (about 15 min), then exit
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PROFILES = os.path.join(ROOT, "profiles")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_info():
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    threads = max(1, min(threads, os.cpu_count() or 1))
    return {"nproc": os.cpu_count(), "cpu_model": model, "threads": threads}


def traffic_from_profile(name):
    """Fabric bytes per launch from a committed counter summary (FETCH_SIZE and
    WRITE_SIZE scaled by the copy-kernel calibration), or None."""
    path = os.path.join(PROFILES, name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def cpu_baseline(A, p, b, dr, dc, V, warm, timed, what):
    """Oracle (restated reference CPU path, u128 OpenMP loop over all the
    threads this host gives the job) on a prefix of the same workload:
    iterations/s from the difference of two runs (set-up and warm-up cancel)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    h = host_info()
    t0 = time.perf_counter()
    O.krylov(A, p, b, dr, dc, V, 0, warm, nthreads=h["threads"])
    t1 = time.perf_counter()
    O.krylov(A, p, b, dr, dc, V, 0, warm + timed, nthreads=h["threads"])
    t2 = time.perf_counter()
    per_iter = ((t2 - t1) - (t1 - t0)) / timed
    return {"value": round(1.0 / per_iter, 4), "unit": "iters/s", "cores": h["threads"], "kind": "port",
            "sample": f"prefix: S_{warm}..S_{warm + timed - 1} of the same {what} after {warm} warm-up entries "
                      f"(oracle/krylov_oracle.c, {h['threads']} OpenMP threads: u128 row-parallel SpMV and "
                      f"pair-parallel projection; set-up cancelled by differencing two runs)",
            "nproc": h["nproc"], "cpu_model": h["cpu_model"]}


def krylov_leg(engine, A, p, b, dr, dc, V, steps, warmup, prof_steps):
    """Timed Krylov iterations on one matrix, then a profiled stretch with
    per-kernel events.  Returns (rate dict, sequence produced before timing)."""
    K = engine.Krylov(A, p, b, dr, dc, V)
    warm = K.run(max(warmup, 1))
    t1 = time.perf_counter()
    K.run(steps)
    wall = time.perf_counter() - t1
    K.set_profiling(True)
    K.run(prof_steps)
    st = K.stats()
    K.close()
    pa, pt = max(st["passes_a"], 1), max(st["passes_at"], 1)
    passes = st["passes_a"] + st["passes_at"]
    spmm_ms = (st["ms_a"] + st["ms_at"]) / max(passes, 1)
    proj_ms = st["ms_proj"] / max(passes, 1)
    bytes_pass = (st["bytes_a"] * st["passes_a"] + st["bytes_at"] * st["passes_at"]) / max(passes, 1)
    log(f"  pass by A {st['ms_a'] / pa * 1e3:.1f} us, by A^T {st['ms_at'] / pt * 1e3:.1f} us, projection "
        f"{proj_ms * 1e3:.1f} us/pass; timed {steps} iterations in {wall * 1e3:.1f} ms")
    return {"wall_s": wall, "iters_s": steps / wall, "spmm_ms": spmm_ms, "proj_ms": proj_ms,
            "a_us": st["ms_a"] / pa * 1e3, "at_us": st["ms_at"] / pt * 1e3, "bytes_pass": bytes_pass,
            "stream_ms": st["ms_total"], "prof_steps": prof_steps,
            "setup": {"ms": round(st["ms_setup"], 1), "on_device": bool(st["setup_on_device"])}}, warm


def roofline(r, traffic, kernel, copy_gbs, gather=None):
    achieved = r["bytes_pass"] / (r["spmm_ms"] * 1e-3) / 1e9
    out = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kernel,
           "bytes_per_launch": int(r["bytes_pass"]), "avg_launch_us": round(r["spmm_ms"] * 1e3, 2),
           "pass_a_us": round(r["a_us"], 2), "pass_at_us": round(r["at_us"], 2)}
    if traffic:
        out["traffic_gbs"] = round(traffic / (r["spmm_ms"] * 1e-3) / 1e9, 1)
        out["traffic_frac"] = round(out["traffic_gbs"] / HBM_PEAK_GBS, 4)
    if r["proj_ms"] > 0:
        out["projection_kernel"] = {"name": "k_proj64", "avg_launch_us": round(r["proj_ms"] * 1e3, 2)}
    if copy_gbs:
        out["copy_ceiling_gbs"] = round(copy_gbs, 1)
    if gather:  # the same number of random 256-byte row gathers, from vectors of the passes' sizes
        for key, runs in gather.items():
            g_us = sum(ms for ms, _ in runs) / len(runs) * 1e3
            out[key] = {"avg_us": round(g_us, 2), "gbs": round(sum(g for _, g in runs) / len(runs), 1),
                        "pass_time_over_probe_time": round(r["spmm_ms"] * 1e3 / g_us, 3)}
    return out


def graph_complex_leg(engine, synth, p, steps=0, with_pm=True, pm_order=0, cpu_timed=0):
    """The 200000 x 180000 graph-complex workload at b = 32 over its whole
    sequence (11258 entries; `steps` timed entries instead if given), its
    device PM-basis of order L - 1 and rank, the same rank through one
    rr_rank call, and (cpu_timed > 0) the oracle's Krylov loop on a prefix of
    the same sequence."""
    b = 32
    A = synth.graph_complex()
    dr, dc, V = synth.preconditioners(2, A.m, A.n, b, p)
    L = synth.seq_len_plain(min(A.m, A.n), b)
    steps = steps or L
    log(f"graph_complex {A.m}x{A.n} nnz={A.nnz} b={b} L={L}")
    K = engine.Krylov(A, p, b, dr, dc, V)
    need = (min(pm_order, L - 1) + 1 if pm_order else L) if with_pm else min(L, 64)
    seq = K.run(need)  # the sequence for the PM-basis; also sizes the device buffer
    t1 = time.perf_counter()
    K.run(steps)  # timed: a whole sequence's worth of iterations in one call
    wall = time.perf_counter() - t1
    K.set_profiling(True)
    K.run(min(steps, 2000))
    st = K.stats()
    K.close()
    passes = st["passes_a"] + st["passes_at"]
    r = {"spmm_ms": (st["ms_a"] + st["ms_at"]) / passes, "proj_ms": 0.0,
         "a_us": st["ms_a"] / st["passes_a"] * 1e3, "at_us": st["ms_at"] / st["passes_at"] * 1e3,
         "bytes_pass": (st["bytes_a"] * st["passes_a"] + st["bytes_at"] * st["passes_at"]) / passes}
    out = {"workload": f"graph_complex {A.m}x{A.n} 12 nnz/row Zipf(1.1) +-1/+-2 p=2013265921 b={b}",
           "seq_len": L, "steps": steps, "iters_s": round(steps / wall, 1),
           "ms_per_step": round(wall * 1e3 / steps, 5),
           "setup": {"ms": round(st["ms_setup"], 1), "on_device": bool(st["setup_on_device"])},
           "roofline": roofline(r, traffic_from_profile("traffic_k_spmm_graph_complex.json"),
                                "k_spmm<32, packed>", None)}
    log(f"  {out['iters_s']} iters/s, pass by A {r['a_us']:.1f} us, by A^T {r['at_us']:.1f} us")
    if not with_pm:
        return out
    d = min(pm_order, L - 1) if pm_order else L - 1  # a shorter order for counter runs
    t3 = time.perf_counter()
    basis, delta = engine.pm_basis(seq[: d + 1], b, d, p)
    t4 = time.perf_counter()
    rank = engine.rank_from_basis(basis, delta, b, p)
    del basis
    out.update({"pm_order": d, "pm_basis_ms": round((t4 - t3) * 1e3, 1), "rank": int(rank)})
    if d == L - 1:  # the same rank through one rr_rank call
        rr_rank, rr_ms = engine.rank(A, p, b, dr, dc, V, L, L - 1)
        assert int(rr_rank) == int(rank), (rr_rank, rank)
        out["rr_rank_s"] = round(rr_ms / 1e3, 3)
    log(f"  PM-basis of order {d}: {out['pm_basis_ms']} ms; rank {rank}; rr_rank {out.get('rr_rank_s')} s")
    if cpu_timed:
        out["cpu_baseline"] = cpu_baseline(A, p, b, dr, dc, V, 2, cpu_timed, "200000x180000 b=32 workload")
        log(f"  CPU baseline {out['cpu_baseline']['value']} iters/s on {out['cpu_baseline']['cores']} threads")
    return out


def plain_small_leg(engine, synth, p, reps=20):
    """The 2000 x 1800 plain workload (1800 base rows of 8 uniform nonzeros,
    200 rows combining 3 base rows each; b = 8, L = 458, order 457) end to end:
    rr_rank on the GPU (best of `reps` calls) against the oracle's whole
    pipeline on this host (Krylov loop, PM-basis, rank formula), both checked
    against dense Gaussian elimination."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    b = 8
    A = synth.plain_small()
    dr, dc, V = synth.preconditioners(1, A.m, A.n, b, p)
    L = synth.seq_len_plain(min(A.m, A.n), b)
    dense = O.dense_rank(A.to_dense(p), p)
    gpu = []
    for _ in range(reps):
        r, ms = engine.rank(A, p, b, dr, dc, V, L, L - 1)
        assert int(r) == dense, (r, dense)
        gpu.append(ms)
    h = host_info()
    t0 = time.perf_counter()
    seq = O.krylov(A, p, b, dr, dc, V, 0, L, nthreads=h["threads"])
    basis, delta = O.pm_basis(seq, b, L - 1, p)
    r_cpu = O.rank_from_basis(basis, delta, b, p)
    t_cpu = time.perf_counter() - t0
    assert int(r_cpu) == dense, (r_cpu, dense)
    out = {"workload": f"plain_small {A.m}x{A.n} ({A.nnz} nonzeros), p=2013265921 b={b}", "seq_len": L,
           "order": L - 1, "rank": dense, "dense_ge_rank": dense, "gpu_rr_rank_ms": round(min(gpu), 2),
           "gpu_rr_rank_ms_median": round(float(np.median(gpu)), 2), "cpu_oracle_ms": round(t_cpu * 1e3, 1),
           "cpu_threads": h["threads"]}
    log(f"plain_small: rank {dense} (dense GE); rr_rank {out['gpu_rr_rank_ms']} ms, oracle pipeline "
        f"{out['cpu_oracle_ms']} ms on {h['threads']} threads")
    return out


def batched(eng, synth, p, nmat, seed=4, plain=True):
    """Batched form: nmat matrices of the batched workload (rows log-uniform in
    [1000, 50000], cols 0.9 rows, 6-16 nnz/row, +-1/+-2, 1-5 % dependent
    columns), b per matrix from the autotuner, solved in one rr_batch_rank
    (timed after one untimed call on the same matrices), vs the same matrices
    solved one at a time by the plain GPU path (rr_rank)."""
    t0 = time.perf_counter()
    mats = synth.batch_matrices(seed, nmat)
    bs, sus, pred = eng.autotune(mats, p)
    _, _, probe = eng.autotune_b(mats[:1], p)
    pre = [synth.preconditioners(5000 + i, A.m, A.n, int(b), p) for i, (A, b) in enumerate(zip(mats, bs))]
    hist = lambda a: {int(k): int(v) for k, v in zip(*np.unique(a, return_counts=True))}
    log(f"batched: {nmat} matrices, rows {min(A.m for A in mats)}..{max(A.m for A in mats)}, "
        f"b histogram {hist(bs)}, su histogram {hist(sus)}, setup {time.perf_counter() - t0:.1f}s")
    # untimed warm-up call on the same matrices: kernels loaded and the device
    # memory cache holding blocks of this workload's sizes, as the plain solves
    # after it find them (they reuse the cache the batched call leaves)
    tw = time.perf_counter()
    eng.batch_rank(mats, p, bs, [x[0] for x in pre], [x[1] for x in pre], [x[2] for x in pre], sus=sus)
    log(f"  warm-up batched call {time.perf_counter() - tw:.2f}s")
    t2 = time.perf_counter()
    ranks, st = eng.batch_rank(mats, p, bs, [x[0] for x in pre], [x[1] for x in pre], [x[2] for x in pre], sus=sus)
    t_batch = time.perf_counter() - t2
    out = {"workload": "batched: rows log-uniform in [1000, 50000], cols 0.9 rows, 6-16 nnz/row, +-1/+-2, "
                       "1-5 % planted dependent columns", "matrices": nmat,
           "ranks_per_s": round(nmat / t_batch, 3), "wall_s": round(t_batch, 3), "ticks": st["ticks"],
           "pm_rounds": st["pm_rounds"], "ms_setup": round(st["ms_setup"], 1), "ms_krylov": round(st["ms_krylov"], 1),
           "ms_pm": round(st["ms_pm"], 1),
           "b_hist": hist(bs), "su_hist": hist(sus), "autotune_probe": probe}
    if not plain:  # the batched call alone (for kernel traces)
        log(f"batched {t_batch:.2f}s (plain solves skipped)")
        return out
    plain_rank = []
    t1 = time.perf_counter()
    for (A, b, (dr, dc, V)) in zip(mats, bs, pre):
        plain_rank.append(eng.rank(A, p, int(b), dr, dc, V)[0])
    t_plain = time.perf_counter() - t1
    assert list(ranks) == plain_rank, "batched ranks differ from plain ranks"
    out.update({"plain_ranks_per_s": round(nmat / t_plain, 3), "speedup_vs_plain": round(t_plain / t_batch, 2),
                "plain_wall_s": round(t_plain, 3), "ranks_equal_plain": True})
    log(f"batched {t_batch:.2f}s vs plain {t_plain:.2f}s -> {t_plain / t_batch:.1f}x")
    return out


def _serve(eng, p, reqs, arrivals):
    """Submits reqs (A, b, dr, dc, V) to a fresh rr_server at the given arrival
    times (s from the start); returns (ranks, latencies ms, wall s, stats)."""
    import threading
    n = len(reqs)
    results = {}
    srv = eng.Server(p, 200 << 30)
    done = threading.Event()

    def poller():
        while len(results) < n:
            for r in srv.poll(wait_ms=50):
                results[r["id"]] = r
                if len(results) % 16 == 0:
                    log(f"  served {len(results)}/{n} ({time.perf_counter() - t0:.1f} s)")
        done.set()

    t0 = time.perf_counter()
    th = threading.Thread(target=poller)
    th.start()
    ids = []
    for (A, b, dr, dc, V), a in zip(reqs, arrivals):
        delay = t0 + a - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        ids.append(srv.submit(A, int(b), dr, dc, V))
    done.wait(timeout=900)
    wall = time.perf_counter() - t0
    th.join(timeout=5)
    st = srv.stats()
    srv.close()
    ranks = [results[i]["rank"] for i in ids]
    lat = np.array([results[i]["latency_ms"] for i in ids])
    assert all(results[i]["status"] == 0 for i in ids), "a served request failed"
    return ranks, lat, wall, st


def served(eng, synth, p, nreq, rate_frac, rows_hi, seed=5):
    """Served form: nreq requests of the served workload (rows log-uniform in
    [10^3, rows_hi], 8-20 nnz/row, +-1/+-2, planted deficits), b per request
    from the autotuner, through rr_server with a 200 GB HBM budget.  Two
    points: all requests at once (saturated: the server's capacity), and
    Poisson arrivals at rate_frac x that capacity, the latter replayed as FIFO
    plain solves (rr_rank one at a time in arrival order, run for real)."""
    t0 = time.perf_counter()
    mats = synth.served_matrices(seed, nreq, rows_hi)
    bs, _, _ = eng.autotune_b(mats, p)
    reqs = [(A, int(b)) + tuple(synth.preconditioners(9000 + i, A.m, A.n, int(b), p))
            for i, (A, b) in enumerate(zip(mats, bs))]
    log(f"served: {nreq} requests, rows {min(A.m for A in mats)}..{max(A.m for A in mats)}, "
        f"b histogram {dict(zip(*np.unique(bs, return_counts=True)))}, setup {time.perf_counter() - t0:.1f}s")
    # saturated: every request at t = 0
    r_sat, lat_sat, wall_sat, st_sat = _serve(eng, p, reqs, np.zeros(nreq))
    cap = nreq / wall_sat
    # offered load rate_frac x capacity, Poisson
    lam = rate_frac * cap
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / lam, nreq))
    r_off, lat_off, wall_off, st_off = _serve(eng, p, reqs, arrivals)
    # FIFO plain on the same arrivals, run: one solve at a time in arrival order
    t1 = time.perf_counter()
    fifo, plain_rank, service = [], [], 0.0
    for (A, b, dr, dc, V), a in zip(reqs, arrivals):
        delay = t1 + a - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        s0 = time.perf_counter()
        plain_rank.append(eng.rank(A, p, b, dr, dc, V)[0])
        fin = time.perf_counter()
        service += fin - s0
        fifo.append((fin - t1 - a) * 1e3)
        if len(fifo) % 16 == 0:
            log(f"  FIFO plain {len(fifo)}/{nreq} ({fin - t1:.1f} s)")
    fifo = np.array(fifo)
    wall_fifo = time.perf_counter() - t1
    assert r_sat == plain_rank and r_off == plain_rank, "served ranks differ from plain ranks"
    pct = lambda x, q: round(float(np.percentile(x, q)), 2)
    out = {"workload": f"served (rehearsal): rows log-uniform in [1000, {rows_hi}], cols 0.9 rows, 8-20 nnz/row, "
                       f"+-1/+-2, 1-5 % planted dependent columns; HBM budget 200 GB", "requests": nreq,
           "saturated": {"ranks_per_s": round(cap, 3), "p50_ms": pct(lat_sat, 50), "p99_ms": pct(lat_sat, 99),
                         "max_active": st_sat["max_active"]},
           "offered": {"rate_frac": rate_frac, "offered_rate_per_s": round(lam, 3),
                       "ranks_per_s": round(nreq / wall_off, 3), "p50_ms": pct(lat_off, 50),
                       "p99_ms": pct(lat_off, 99), "max_active": st_off["max_active"]},
           "fifo_plain": {"ranks_per_s": round(nreq / wall_fifo, 3), "p50_ms": pct(fifo, 50),
                          "p99_ms": pct(fifo, 99), "service_ranks_per_s": round(nreq / service, 3)},
           "saturated_vs_plain": round(cap * service / nreq, 2), "ranks_equal_plain": True}
    log(f"served: saturated {cap:.1f} ranks/s (p50 {out['saturated']['p50_ms']} ms); at {rate_frac} x: p50 "
        f"{out['offered']['p50_ms']} p99 {out['offered']['p99_ms']} ms; FIFO plain p50 {out['fifo_plain']['p50_ms']} "
        f"p99 {out['fifo_plain']['p99_ms']} ms; plain service {nreq / service:.1f} ranks/s "
        f"-> {out['saturated_vs_plain']}x")
    return out


def full_rank(engine, synth, p):
    """The large single matrix end to end through rr_rank: 140633 Krylov
    entries at b = 64 and the PM-basis of order 140632 beside them; the rank
    must be the planted n - 5000 = 4495000."""
    b = 64
    t0 = time.perf_counter()
    A = synth.large_single()
    dr, dc, V = synth.preconditioners(2, A.m, A.n, b, p)
    L = synth.seq_len_plain(min(A.m, A.n), b)
    log(f"large_single {A.m}x{A.n} nnz={A.nnz} b={b} L={L}, generated in {time.perf_counter() - t0:.1f}s")
    os.environ.setdefault("RR_PROGRESS", "20")  # the library's own progress lines on stderr
    rank, ms = engine.rank(A, p, b, dr, dc, V, L, L - 1)
    expect = A.n - 5000
    out = {"workload": f"large_single {A.m}x{A.n} 20 nnz/row uniform, values uniform in [1,p), 5000 planted "
                       f"dependent columns, p=2013265921 b={b}", "seq_len": L, "order": L - 1, "rank": int(rank),
           "planted_rank": expect, "rank_ok": int(rank) == expect, "rr_rank_s": round(ms / 1e3, 2),
           "nnz": A.nnz}
    print(json.dumps(out), flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="timed iterations (default 200)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-timed", type=int, default=16, help="timed oracle iterations for the CPU baseline")
    ap.add_argument("--cpu-gc-timed", type=int, default=256,
                    help="timed oracle iterations for the graph-complex CPU baseline")
    ap.add_argument("--no-graph-complex", action="store_true", help="skip the 200000 x 180000 workload")
    ap.add_argument("--no-pm", action="store_true", help="graph_complex: skip its PM-basis and rank")
    ap.add_argument("--pm-order", type=int, default=0,
                    help="graph_complex: PM-basis of this order on the sequence's first entries (counter runs)")
    ap.add_argument("--batched", type=int, default=512, help="matrices in the batched comparison (0: skip)")
    ap.add_argument("--served", type=int, default=256, help="requests in the served comparison (0: skip)")
    ap.add_argument("--served-rows-hi", type=int, default=100000,
                    help="largest rows of a served request (the full workload: 1000000)")
    ap.add_argument("--rate-frac", type=float, default=0.7, help="served offered load / saturated capacity")
    ap.add_argument("--no-plain", action="store_true",
                    help="batched call only: skip the plain solves and the served form (for kernel traces)")
    ap.add_argument("--workload", choices=["large", "graph_complex", "batched", "served"], default="large",
                    help="graph_complex / batched: only that leg (for kernel traces and counters)")
    ap.add_argument("--full-rank", action="store_true", help="only the large matrix's rank end to end")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MB",
                    help="only run the plain copy kernel on MB megabytes (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}): the engine does not shard a matrix "
              f"across GPUs, and runs on one", file=sys.stderr)
        sys.exit(2)
    from rust_rank_amd import engine, synth
    p = synth.P_DEFAULT
    if args.copy_calib:
        ms, gbs = engine.copy_probe(args.copy_calib << 20, 20)
        print(json.dumps({"copy_bytes": args.copy_calib << 20, "avg_ms": ms, "read_write_gbs": gbs}), flush=True)
        return
    if args.full_rank:
        full_rank(engine, synth, p)
        return
    if args.workload == "graph_complex":
        print(json.dumps(graph_complex_leg(engine, synth, p, args.steps, not args.no_pm, args.pm_order)), flush=True)
        return
    if args.workload == "batched":
        print(json.dumps(batched(engine, synth, p, args.batched or 512, plain=not args.no_plain)), flush=True)
        return
    if args.workload == "served":
        print(json.dumps(served(engine, synth, p, args.served or 256, args.rate_frac, args.served_rows_hi)), flush=True)
        return

    b = 64
    steps = args.steps or 200
    t0 = time.perf_counter()
    A = synth.large_single()
    dr, dc, V = synth.preconditioners(2, A.m, A.n, b, p)
    L = synth.seq_len_plain(min(A.m, A.n), b)
    desc = (f"large_single {A.m}x{A.n} 20 nnz/row uniform, values uniform in [1,p), 5000 planted dependent "
            f"columns, p=2013265921 b={b}")
    log(f"workload {desc}: nnz={A.nnz} L={L}; generated in {time.perf_counter() - t0:.1f}s")
    r, _ = krylov_leg(engine, A, p, b, dr, dc, V, steps, args.warmup, 40)
    _, copy_gbs = engine.copy_probe(1 << 30, 20)
    # by A gathers rows of x (n rows), by A^T of y (m rows): the gathers alone, then with the pass's
    # value reads and output rows
    gather = {"gather_probe": [engine.gather_probe(rows, A.nnz, 5) for rows in (A.n, A.m)],
              "pass_traffic_probe": [engine.gather_probe(rows, A.nnz, 5, True) for rows in (A.n, A.m)]}
    b_iter = 8 * A.nnz + (4 * b + 2) * (A.m + A.n)
    roof = roofline(r, traffic_from_profile("traffic_k_spmm_large_single.json"), "k_spmm<64, general>", copy_gbs,
                    gather)
    cpu = None if args.no_cpu else cpu_baseline(A, p, b, dr, dc, V, 2, args.cpu_timed,
                                                "5000000x4500000 b=64 workload")
    nnz = A.nnz
    del A
    gc = None if args.no_graph_complex else graph_complex_leg(engine, synth, p,
                                                              cpu_timed=0 if args.no_cpu else args.cpu_gc_timed)
    small = plain_small_leg(engine, synth, p)
    bat = srv = None
    engine.release_device_memory()  # the large legs' cached blocks go back before the many small matrices
    if args.batched:
        bat = batched(engine, synth, p, args.batched, plain=not args.no_plain)
    if args.served and not args.no_plain:
        srv = served(engine, synth, p, args.served, args.rate_frac, args.served_rows_hi)
    out = {
        "metric": "Krylov iters/s + HBM GB/s; batched ranks/s & p50/p99 vs plain, same GPU",
        "value": round(r["iters_s"], 3), "unit": "iters/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": round(r["wall_s"] * 1e3 / steps, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32", "data": "synthetic (seeded SplitMix64)",
        "config": {"workload": desc, "seq_len": L, "nnz": nnz, "b": b,
                   "hbm_gbs_algorithmic": round(b_iter * r["iters_s"] / 1e9, 1)},
        "setup": r["setup"],
        "roofline": roof, "cpu_baseline": cpu, "graph_complex": gc, "plain_small": small, "batched": bat,
        "served": srv,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
