"""CPU-side checks of the C ABI: the library loads, exports every function
include/rank_engine.h declares, validates arguments before touching a GPU, and
its host-side logic (length policy, rank formula) matches the reference."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from rust_rank_amd import engine, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "rank_engine.h")
P = 2013265921


def declared_functions():
    text = open(HDR).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rr_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 15
    lib = C.CDLL(engine.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_invalid_arguments_fail_before_device_work():
    A = synth.uniform_sparse(1, 1, 20, 10, 1, 3, P, small=False)
    dr, dc, V = synth.preconditioners(1, A.m, A.n, 4, P)
    with pytest.raises(engine.RankEngineError) as e:
        engine.Krylov(A, P, 0, dr, dc, V)
    assert e.value.status == -1
    with pytest.raises(engine.RankEngineError):
        engine.Krylov(A, 15, 4, dr, dc, V)  # not prime
    with pytest.raises(engine.RankEngineError):
        engine.pm_basis(np.zeros((5, 10), np.uint32), 4, 5, P)  # d >= L


def test_length_and_order_policies():
    # exact policy: 2 min(m, n) / b + 8 entries, order L - 1
    assert engine.seq_len(2000, 1800, 8) == 458
    assert engine.pm_order(458) == 457
    assert engine.seq_len(200000, 180000, 32) == 11258
    # reference policy (src/main.rs:497-504, 599-603)
    assert engine.seq_len(2000, 1800, 8, True) == 512 + 1
    assert engine.pm_order(513, True) == 512
    assert engine.seq_len(100, 90, 1, True) == int(2 * 90 / 1 + 4) + 1


def test_rank_formula_matches_reference_golden():
    d = np.load(os.path.join(ROOT, "tests", "golden", "pm_basis.npz"))
    i = 0
    while f"seq{i}" in d:
        p, b, dd, L = (int(x) for x in d[f"meta{i}"])
        assert engine.rank_from_basis(d[f"basis{i}"], d[f"delta{i}"], b, p) == int(d[f"rank{i}"][0]), i
        i += 1


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus above 1 exits with status 2 and says
    why, before any device work."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""
