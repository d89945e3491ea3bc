"""ctypes binding of librank_engine.so (include/rank_engine.h).

The library is the product: there is no CPU fallback.  Importing this module
on a machine without the built library raises, and every call that fails on
the device raises RankEngineError with the library's message.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librank_engine.so")

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


class RankEngineError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"rank engine status {status}: {msg}")
        self.status = status


class _Csr(C.Structure):
    _fields_ = [("m", C.c_uint64), ("n", C.c_uint64), ("row_ptr", C.c_void_p), ("col", C.c_void_p),
                ("val", C.c_void_p)]


class KrylovStats(C.Structure):
    _fields_ = [("passes_a", C.c_uint64), ("passes_at", C.c_uint64), ("ms_a", C.c_double), ("ms_at", C.c_double),
                ("ms_reduce", C.c_double), ("ms_total", C.c_double),
                ("bytes_a", C.c_uint64), ("bytes_at", C.c_uint64), ("ms_proj", C.c_double),
                ("ms_setup", C.c_double), ("setup_on_device", C.c_int32), ("reserved", C.c_int32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class BatchStats(C.Structure):
    _fields_ = [("ms_setup", C.c_double), ("ms_krylov", C.c_double), ("ms_pm", C.c_double), ("ms_total", C.c_double),
                ("ticks", C.c_uint64), ("pm_rounds", C.c_uint64), ("pm_launches", C.c_uint64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Result(C.Structure):
    _fields_ = [("id", C.c_uint64), ("rank", C.c_int64), ("status", C.c_int32), ("latency_ms", C.c_double),
                ("queue_ms", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("queued", C.c_uint64), ("used_bytes", C.c_uint64), ("completed", C.c_uint64),
                ("failed", C.c_uint64), ("ticks", C.c_uint64), ("pm_rounds", C.c_uint64), ("max_active", C.c_uint64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None


def lib():
    """Loads librank_engine.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m rust_rank_amd.build`")
    L = C.CDLL(LIB_PATH)
    L.rr_last_error.restype = C.c_char_p
    L.rr_device_count.restype = C.c_int
    L.rr_release_device_memory.argtypes = []
    L.rr_krylov_create.argtypes = [C.POINTER(_Csr), C.c_uint32, C.c_uint32, _u32p, _u32p, _u32p, C.c_uint64,
                                   C.POINTER(C.c_void_p)]
    L.rr_krylov_run.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p]
    L.rr_krylov_state.argtypes = [C.c_void_p, _u32p, C.POINTER(C.c_uint64)]
    L.rr_krylov_set_profiling.argtypes = [C.c_void_p, C.c_int]
    L.rr_krylov_stats_get.argtypes = [C.c_void_p, C.POINTER(KrylovStats)]
    L.rr_krylov_destroy.argtypes = [C.c_void_p]
    L.rr_krylov_destroy.restype = None
    L.rr_pm_basis.argtypes = [_u32p, C.c_uint32, C.c_uint64, C.c_uint64, C.c_uint32, _u32p, _i64p]
    L.rr_rank_from_basis.argtypes = [_u32p, _i64p, C.c_uint32, C.c_uint64, C.c_uint32, C.POINTER(C.c_int64)]
    L.rr_ntt.argtypes = [_u32p, C.c_uint32, C.c_uint32, C.c_int]
    L.rr_poly_mat_mul.argtypes = [_u32p, C.c_uint64, _u32p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32,
                                  C.c_uint64, C.c_uint64, C.c_uint32, _u32p]
    L.rr_seq_len.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_int]
    L.rr_seq_len.restype = C.c_uint64
    L.rr_pm_order.argtypes = [C.c_uint64, C.c_int]
    L.rr_pm_order.restype = C.c_uint64
    L.rr_rank.argtypes = [C.POINTER(_Csr), C.c_uint32, C.c_uint32, _u32p, _u32p, _u32p, C.c_uint64, C.c_uint64,
                          C.POINTER(C.c_int64), C.POINTER(C.c_double)]
    L.rr_rank.restype = C.c_int
    L.rr_batch_rank.argtypes = [C.c_uint32, C.c_void_p, C.c_uint32, _u32p, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_void_p, C.c_void_p, _i64p, C.POINTER(BatchStats)]
    L.rr_batch_rank.restype = C.c_int
    L.rr_server_create.argtypes = [C.c_uint32, C.c_uint64, C.POINTER(C.c_void_p)]
    L.rr_server_submit.argtypes = [C.c_void_p, C.POINTER(_Csr), C.c_uint32, _u32p, _u32p, _u32p, C.c_uint64,
                                   C.c_uint64, C.POINTER(C.c_uint64)]
    L.rr_server_poll.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32), C.c_int]
    L.rr_server_stats_get.argtypes = [C.c_void_p, C.POINTER(ServerStats)]
    L.rr_server_destroy.argtypes = [C.c_void_p]
    L.rr_server_destroy.restype = None
    for name in ("rr_server_create", "rr_server_submit", "rr_server_poll", "rr_server_stats_get"):
        getattr(L, name).restype = C.c_int
    L.rr_batch_rank_tuned.argtypes = [C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p, _i64p, C.POINTER(BatchStats)]
    L.rr_autotune_b.argtypes = [C.c_uint32, C.c_void_p, C.c_uint32, _u32p, C.c_void_p]
    L.rr_autotune_probe.argtypes = [C.c_void_p, C.c_void_p]
    L.rr_autotune.argtypes = [C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p]
    for name in ("rr_autotune_b", "rr_autotune_probe", "rr_autotune", "rr_batch_rank_tuned"):
        getattr(L, name).restype = C.c_int
    L.rr_copy_probe.argtypes = [C.c_uint64, C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.rr_copy_probe.restype = C.c_int
    L.rr_gather_probe.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_int, C.POINTER(C.c_double),
                                  C.POINTER(C.c_double)]
    L.rr_gather_probe.restype = C.c_int
    _u64ptr = C.POINTER(C.c_uint64)
    L.rr_sms_open.argtypes = [C.c_char_p, C.c_uint32, C.POINTER(C.c_void_p), _u64ptr, _u64ptr, _u64ptr]
    L.rr_sms_copy.argtypes = [C.c_void_p, _u64p, _u32p, _u32p]
    L.rr_sms_close.argtypes = [C.c_void_p]
    L.rr_sms_close.restype = None
    L.rr_wdm_info.argtypes = [C.c_char_p, _u64ptr, _u64ptr, C.POINTER(C.c_uint32), _u64ptr, C.POINTER(C.c_uint32)]
    L.rr_wdm_load.argtypes = [C.c_char_p, _u32p, _u32p, _u32p, _u32p, _u32p]
    L.rr_rank_from_wdm.argtypes = [C.c_char_p, C.c_int, C.c_uint64, C.POINTER(C.c_int64)]
    L.rr_wdm_save.argtypes = [C.c_char_p, C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, _u32p, _u32p, _u32p,
                              _u32p, _u32p, C.c_uint64]
    L.rr_seq_resume_count.argtypes = [C.c_uint64, C.c_uint64]
    L.rr_seq_resume_count.restype = C.c_uint64
    for name in ("rr_sms_open", "rr_sms_copy", "rr_wdm_info", "rr_wdm_load", "rr_wdm_save"):
        getattr(L, name).restype = C.c_int
    for name in ("rr_krylov_create", "rr_krylov_run", "rr_krylov_state", "rr_krylov_set_profiling",
                 "rr_krylov_stats_get", "rr_pm_basis", "rr_rank_from_basis", "rr_ntt", "rr_poly_mat_mul"):
        getattr(L, name).restype = C.c_int
    _lib = L
    return L


def _check(status: int):
    if status != 0:
        raise RankEngineError(status, lib().rr_last_error().decode())


def device_count() -> int:
    return int(lib().rr_device_count())


def release_device_memory():
    """Returns the library's cached device buffers to the system."""
    _check(lib().rr_release_device_memory())


def _csr_struct(A):
    rp = np.ascontiguousarray(A.row_ptr, dtype=np.uint64)
    col = np.ascontiguousarray(A.col, dtype=np.uint32)
    val = np.ascontiguousarray(A.val, dtype=np.uint32)
    s = _Csr(A.m, A.n, rp.ctypes.data, col.ctypes.data, val.ctypes.data)
    return s, (rp, col, val)


class Krylov:
    """Device Krylov stream over B = D_c A^T D_r A D_c (rr_krylov_*)."""

    def __init__(self, A, p: int, b: int, row_precond, col_precond, curv, t_have: int = 0):
        s, keep = _csr_struct(A)
        self.b, self.p, self.n = b, p, A.n
        self.P = b * (b + 1) // 2
        h = C.c_void_p()
        _check(lib().rr_krylov_create(C.byref(s), p, b, np.ascontiguousarray(row_precond, dtype=np.uint32),
                                      np.ascontiguousarray(col_precond, dtype=np.uint32),
                                      np.ascontiguousarray(curv, dtype=np.uint32), t_have, C.byref(h)))
        self._h = h

    def run(self, count: int) -> np.ndarray:
        out = np.zeros((max(count, 1), self.P), np.uint32)
        _check(lib().rr_krylov_run(self._h, count, out.ctypes.data))
        return out[:count]

    def state(self):
        curv = np.zeros((self.n, self.b), np.uint32)
        t = C.c_uint64()
        _check(lib().rr_krylov_state(self._h, curv, C.byref(t)))
        return curv, t.value

    def set_profiling(self, on: bool = True):
        _check(lib().rr_krylov_set_profiling(self._h, int(on)))

    def stats(self) -> dict:
        st = KrylovStats()
        _check(lib().rr_krylov_stats_get(self._h, C.byref(st)))
        return st.as_dict()

    def close(self):
        if getattr(self, "_h", None):
            lib().rr_krylov_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def pm_basis(seq, b: int, d: int, p: int):
    """Order-d PM-basis of [S(x); I_b] on the device (rr_pm_basis)."""
    seq = np.ascontiguousarray(seq, dtype=np.uint32)
    basis = np.zeros((2 * b, 2 * b, d + 1), np.uint32)
    delta = np.zeros(2 * b, np.int64)
    _check(lib().rr_pm_basis(seq, b, seq.shape[0], d, p, basis, delta))
    return basis, delta


def rank_from_basis(basis, delta, b: int, p: int) -> int:
    basis = np.ascontiguousarray(basis, dtype=np.uint32)
    r = C.c_int64()
    _check(lib().rr_rank_from_basis(basis, np.ascontiguousarray(delta, dtype=np.int64), b, basis.shape[2], p,
                                    C.byref(r)))
    return r.value


def ntt(data, inverse: bool = False):
    """Natural-order NTT mod 2013265921 of each row of `data` (rows of 2^logn points)."""
    a = np.array(data, dtype=np.uint32, copy=True, order="C")
    if a.ndim == 1:
        a = a[None, :]
    n = a.shape[1]
    logn = n.bit_length() - 1
    _check(lib().rr_ntt(a, logn, a.shape[0], int(inverse)))
    return a


def poly_mat_mul(A, B, lo: int, hi: int, p: int):
    A = np.ascontiguousarray(A, dtype=np.uint32)
    B = np.ascontiguousarray(B, dtype=np.uint32)
    rows, inner, la = A.shape
    inner2, cols, lb = B.shape
    assert inner == inner2
    out = np.zeros((rows, cols, hi - lo), np.uint32)
    _check(lib().rr_poly_mat_mul(A, la, B, lb, rows, inner, cols, lo, hi, p, out))
    return out


def load_sms(path: str, p: int):
    """SMS text file -> Csr (rr_sms_open / rr_sms_copy), values as residues mod p."""
    from .synth import Csr
    h = C.c_void_p()
    m, n, nnz = C.c_uint64(), C.c_uint64(), C.c_uint64()
    _check(lib().rr_sms_open(os.fsencode(path), p, C.byref(h), C.byref(m), C.byref(n), C.byref(nnz)))
    try:
        rp = np.zeros(m.value + 1, np.uint64)
        col = np.zeros(max(nnz.value, 1), np.uint32)
        val = np.zeros(max(nnz.value, 1), np.uint32)
        _check(lib().rr_sms_copy(h, rp, col, val))
    finally:
        lib().rr_sms_close(h)
    return Csr(m.value, n.value, rp, col[: nnz.value], val[: nnz.value])


def wdm_info(path: str) -> dict:
    m, n, ln = C.c_uint64(), C.c_uint64(), C.c_uint64()
    p, b = C.c_uint32(), C.c_uint32()
    _check(lib().rr_wdm_info(os.fsencode(path), C.byref(m), C.byref(n), C.byref(p), C.byref(ln), C.byref(b)))
    return {"m": m.value, "n": n.value, "p": p.value, "len": ln.value, "b": b.value}


def wdm_load(path: str) -> dict:
    """.wdm checkpoint -> dict(m, n, p, b, row_precond, col_precond, V (n x b), curv (n x b), seq [len][P])."""
    h = wdm_info(path)
    m, n, b, ln = h["m"], h["n"], h["b"], h["len"]
    P = b * (b + 1) // 2
    h.update(row_precond=np.zeros(m, np.uint32), col_precond=np.zeros(n, np.uint32),
             V=np.zeros((n, b), np.uint32), curv=np.zeros((n, b), np.uint32), seq=np.zeros((max(ln, 1), P), np.uint32))
    _check(lib().rr_wdm_load(os.fsencode(path), h["row_precond"], h["col_precond"], h["V"], h["curv"], h["seq"]))
    h["seq"] = h["seq"][:ln]
    return h


def rank_from_wdm(path: str, reference_policy: bool = False, order: int = 0) -> int:
    """Rank from the sequence stored in a .wdm file (rr_rank_from_wdm)."""
    r = C.c_int64()
    _check(lib().rr_rank_from_wdm(os.fsencode(path), int(reference_policy), order, C.byref(r)))
    return r.value


def wdm_save(path: str, m: int, n: int, p: int, row_precond, col_precond, V, curv, seq):
    """Writes a .wdm checkpoint (rr_wdm_save); V, curv n x b, seq [len][b(b+1)/2]."""
    V = np.ascontiguousarray(V, dtype=np.uint32)
    b = V.shape[1]
    seq = np.ascontiguousarray(seq, dtype=np.uint32).reshape(-1, b * (b + 1) // 2)
    _check(lib().rr_wdm_save(os.fsencode(path), m, n, p, b, np.ascontiguousarray(row_precond, dtype=np.uint32),
                             np.ascontiguousarray(col_precond, dtype=np.uint32), V,
                             np.ascontiguousarray(curv, dtype=np.uint32),
                             seq if seq.size else np.zeros((1, 1), np.uint32), seq.shape[0]))


def seq_resume_count(have: int, max_nlen: int) -> int:
    return int(lib().rr_seq_resume_count(have, max_nlen))


def seq_len(m: int, n: int, b: int, reference_policy: bool = False) -> int:
    return int(lib().rr_seq_len(m, n, b, int(reference_policy)))


def pm_order(L: int, reference_policy: bool = False) -> int:
    return int(lib().rr_pm_order(L, int(reference_policy)))


def rank(A, p: int, b: int, row_precond, col_precond, V, seq_len: int = 0, order: int = 0):
    """Full device pipeline (rr_rank): returns (rank, milliseconds)."""
    s, keep = _csr_struct(A)
    r = C.c_int64()
    ms = C.c_double()
    _check(lib().rr_rank(C.byref(s), p, b, np.ascontiguousarray(row_precond, dtype=np.uint32),
                         np.ascontiguousarray(col_precond, dtype=np.uint32), np.ascontiguousarray(V, dtype=np.uint32),
                         seq_len, order, C.byref(r), C.byref(ms)))
    return r.value, ms.value


class Tuning(C.Structure):
    _fields_ = [("b", C.c_uint32), ("su", C.c_uint32), ("pred_ms", C.c_double)]


def batch_rank(mats, p: int, bs, drs, dcs, Vs, seq_lens=None, orders=None, sus=None):
    """Ranks of many matrices in one batched run (rr_batch_rank; with `sus`,
    the column block size per matrix, rr_batch_rank_tuned).  Returns (ranks, stats)."""
    n = len(mats)
    keep = []
    csrs = (_Csr * n)()
    for i, A in enumerate(mats):
        s, k = _csr_struct(A)
        csrs[i] = s
        keep.append(k)

    def ptrs(arrs):
        a = [np.ascontiguousarray(x, dtype=np.uint32) for x in arrs]
        keep.append(a)
        return (C.c_void_p * n)(*[x.ctypes.data for x in a])

    b = np.ascontiguousarray(bs, dtype=np.uint32)
    L = np.ascontiguousarray(seq_lens if seq_lens is not None else np.zeros(n), dtype=np.uint64)
    d = np.ascontiguousarray(orders if orders is not None else np.zeros(n), dtype=np.uint64)
    ranks = np.zeros(n, np.int64)
    st = BatchStats()
    if sus is None:
        _check(lib().rr_batch_rank(n, C.cast(csrs, C.c_void_p), p, b, ptrs(drs), ptrs(dcs), ptrs(Vs),
                                   L.ctypes.data, d.ctypes.data, ranks, C.byref(st)))
    else:
        tune = (Tuning * n)(*[Tuning(int(b[i]), int(sus[i]), 0.0) for i in range(n)])
        _check(lib().rr_batch_rank_tuned(n, C.cast(csrs, C.c_void_p), p, C.cast(tune, C.c_void_p), ptrs(drs),
                                         ptrs(dcs), ptrs(Vs), L.ctypes.data, d.ctypes.data, ranks, C.byref(st)))
    return ranks, st.as_dict()


class Server:
    """Continuous-batching rank server (rr_server_*)."""

    def __init__(self, p: int, hbm_budget_bytes: int):
        h = C.c_void_p()
        _check(lib().rr_server_create(p, hbm_budget_bytes, C.byref(h)))
        self._h = h

    def submit(self, A, b: int, row_precond, col_precond, V, seq_len: int = 0, order: int = 0) -> int:
        s, keep = _csr_struct(A)
        rid = C.c_uint64()
        _check(lib().rr_server_submit(self._h, C.byref(s), b, np.ascontiguousarray(row_precond, dtype=np.uint32),
                                      np.ascontiguousarray(col_precond, dtype=np.uint32),
                                      np.ascontiguousarray(V, dtype=np.uint32), seq_len, order, C.byref(rid)))
        return rid.value

    def poll(self, wait_ms: int = 0, cap: int = 256):
        buf = (Result * cap)()
        n = C.c_uint32()
        _check(lib().rr_server_poll(self._h, C.cast(buf, C.c_void_p), cap, C.byref(n), wait_ms))
        return [{"id": buf[i].id, "rank": buf[i].rank, "status": buf[i].status, "latency_ms": buf[i].latency_ms,
                 "queue_ms": buf[i].queue_ms} for i in range(n.value)]

    def stats(self) -> dict:
        st = ServerStats()
        _check(lib().rr_server_stats_get(self._h, C.byref(st)))
        return st.as_dict()

    def close(self):
        if getattr(self, "_h", None):
            lib().rr_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def autotune_b(mats, p: int):
    """Block width per matrix from the measured-probe model (rr_autotune_b)."""
    n = len(mats)
    keep = []
    csrs = (_Csr * n)()
    for i, A in enumerate(mats):
        s, k = _csr_struct(A)
        csrs[i] = s
        keep.append(k)
    out = np.zeros(n, np.uint32)
    pred = np.zeros(n, np.float64)
    _check(lib().rr_autotune_b(n, C.cast(csrs, C.c_void_p), p, out, pred.ctypes.data))
    c = np.zeros(3)
    pm = np.zeros(3)
    _check(lib().rr_autotune_probe(c.ctypes.data, pm.ctypes.data))
    return out, pred, {"c_ms_per_nnz": c.tolist(), "pm_ms_d255": pm.tolist()}


def autotune(mats, p: int):
    """Block width and column block size per matrix (rr_autotune): (b, su, pred_ms) arrays."""
    n = len(mats)
    keep = []
    csrs = (_Csr * n)()
    for i, A in enumerate(mats):
        s, k = _csr_struct(A)
        csrs[i] = s
        keep.append(k)
    out = (Tuning * n)()
    _check(lib().rr_autotune(n, C.cast(csrs, C.c_void_p), p, C.cast(out, C.c_void_p)))
    return (np.array([t.b for t in out], np.uint32), np.array([t.su for t in out], np.uint32),
            np.array([t.pred_ms for t in out]))


def copy_probe(nbytes: int, iters: int = 20):
    """Plain device copy of nbytes: (avg kernel ms, read+write GB/s)."""
    ms, gbs = C.c_double(), C.c_double()
    _check(lib().rr_copy_probe(nbytes, iters, C.byref(ms), C.byref(gbs)))
    return ms.value, gbs.value


def gather_probe(rows: int, gathers: int, iters: int = 5, pass_mix: bool = False):
    """Random 256-byte row gathers (with pass_mix, plus the Krylov pass's value
    reads and output rows): (avg kernel ms, GB/s of the bytes moved)."""
    ms, gbs = C.c_double(), C.c_double()
    _check(lib().rr_gather_probe(rows, gathers, iters, int(pass_mix), C.byref(ms), C.byref(gbs)))
    return ms.value, gbs.value
