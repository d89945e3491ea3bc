"""ctypes bindings for the CPU oracle (oracle/liboracle.so) and, when it was
built, the reference's compiled C++ (oracle/_ref/libref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, and never by the rust_rank_amd package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle oracle)")
        L = C.CDLL(path)
        L.orc_transpose.argtypes = [C.c_uint64, C.c_uint64, _u64p, _u32p, _u32p, _u64p, _u32p, _u32p]
        L.orc_transpose.restype = C.c_int
        L.orc_scale_rows.argtypes = [C.c_uint64, _u64p, _u32p, _u32p, C.c_uint32]
        L.orc_scale_cols.argtypes = [C.c_uint64, _u32p, _u32p, _u32p, C.c_uint32]
        L.orc_spmm.argtypes = [C.c_uint64, _u64p, _u32p, _u32p, C.c_uint32, _u32p, _u32p, C.c_uint32, C.c_int]
        L.orc_krylov.argtypes = [C.c_uint64, C.c_uint64, _u64p, _u32p, _u32p, C.c_uint32, C.c_uint32,
                                 _u32p, _u32p, _u32p, C.c_uint64, C.c_uint64, _u32p, C.c_void_p, C.c_int]
        L.orc_krylov.restype = C.c_int
        L.orc_dense_rank.argtypes = [C.c_uint64, C.c_uint64, _u32p, C.c_uint32, C.c_int]
        L.orc_dense_rank.restype = C.c_uint64
        L.orc_ntt.argtypes = [_u32p, C.c_uint64, C.c_int]
        L.orc_ntt.restype = C.c_int
        L.orc_ntt_generic.argtypes = [_u64p, C.c_uint64, C.c_int, C.c_uint64, C.c_uint64]
        L.orc_ntt_generic.restype = C.c_int
        L.orc_lsp.argtypes = [C.c_uint32, _u32p, C.c_uint32, _u32p, _u32p, _u32p]
        L.orc_lsp.restype = C.c_int
        L.orc_matrix_inverse.argtypes = [C.c_uint32, _u32p, C.c_uint32, _u32p]
        L.orc_matrix_inverse.restype = C.c_int
        L.orc_modular_rank.argtypes = [C.c_uint32, C.c_uint32, _u32p, C.c_uint32]
        L.orc_modular_rank.restype = C.c_uint32
        L.orc_set_use_ntt.argtypes = [C.c_int]
        L.orc_poly_mat_mul.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, _u32p, C.c_uint64, _u32p, C.c_uint64,
                                       C.c_uint64, C.c_uint64, C.c_uint32, _u32p]
        L.orc_m_basis.argtypes = [C.c_uint32, C.c_uint32, _u32p, _i64p, C.c_uint32, _u32p]
        L.orc_pm_basis.argtypes = [_u32p, C.c_uint32, C.c_uint64, C.c_uint64, C.c_uint32, _u32p, _i64p]
        L.orc_pm_basis.restype = C.c_int
        L.orc_rank_from_basis.argtypes = [_u32p, _i64p, C.c_uint32, C.c_uint64, C.c_uint32]
        L.orc_rank_from_basis.restype = C.c_int64
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(os.path.join(_HERE, "_ref", "libref.so"))


def ref():
    global _ref
    if _ref is None:
        path = os.path.join(_HERE, "_ref", "libref.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/_ref/libref.so is not built (needs the reference checkout)")
        R = C.CDLL(path)
        R.ref_ntt_u64.argtypes = [_u64p, C.c_uint64, C.c_int]
        R.ref_poly_mat_mul_u64.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, _u32p, C.c_uint64, _u32p, C.c_uint64,
                                           C.c_uint64, C.c_uint64, C.c_uint32, _u32p]
        R.ref_poly_mat_mul_u64.restype = C.c_int
        R.ref_lsp_u64.argtypes = [C.c_uint32, _u32p, C.c_uint32, _u32p, _u32p, _u32p]
        R.ref_inverse_u64.argtypes = [C.c_uint32, _u32p, C.c_uint32, _u32p]
        R.ref_inverse_u64.restype = C.c_int
        R.ref_modular_rank_u64.argtypes = [C.c_uint32, C.c_uint32, _u32p, C.c_uint32]
        R.ref_modular_rank_u64.restype = C.c_uint64
        R.ref_m_basis_u64.argtypes = [C.c_uint32, C.c_uint32, _u32p, _i64p, C.c_uint32, _u32p]
        R.ref_pm_basis_u64.argtypes = [_u32p, C.c_uint32, C.c_uint64, C.c_uint64, C.c_uint32, _u32p, C.c_uint64, _i64p]
        R.ref_pm_basis_u64.restype = C.c_int64
        R.ref_analyze_pm_basis_u64.argtypes = [_u32p, _i64p, C.c_uint32, C.c_uint64, C.c_uint32]
        R.ref_analyze_pm_basis_u64.restype = C.c_int64
        R.ref_load_sms.argtypes = [C.c_char_p, C.c_uint32, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                   C.POINTER(C.c_int64), C.c_void_p, C.c_void_p, C.c_void_p]
        R.ref_transpose.argtypes = [C.c_int64, C.c_int64, _i64p, _i32p, _u32p, _i64p, _i32p, _u32p]
        R.ref_rescale.argtypes = [C.c_int64, C.c_int64, _i64p, _i32p, _u32p, C.c_void_p, C.c_void_p, C.c_uint32]
        R.ref_save_wdm.argtypes = [C.c_char_p, C.c_int64, C.c_int64, C.c_uint32, C.c_uint32, _u32p, _u32p, _u32p,
                                   _u32p, _u32p, C.c_int64]
        _ref = R
    return _ref


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# ---------------------------------------------------------------- oracle calls
def transpose(m, n, rp, col, val):
    nnz = int(rp[m])
    rpt = np.zeros(n + 1, np.uint64)
    colt = np.zeros(max(nnz, 1), np.uint32)
    valt = np.zeros(max(nnz, 1), np.uint32)
    rc = lib().orc_transpose(m, n, _c(rp, np.uint64), _c(col, np.uint32) if nnz else np.zeros(1, np.uint32),
                             _c(val, np.uint32) if nnz else np.zeros(1, np.uint32), rpt, colt, valt)
    assert rc == 0
    return rpt, colt[:nnz], valt[:nnz]


def spmm(m, rp, col, val, x, p, nthreads=0):
    x = _c(x, np.uint32)
    b = x.shape[1]
    y = np.zeros((m, b), np.uint32)
    lib().orc_spmm(m, _c(rp, np.uint64), _c(col, np.uint32), _c(val, np.uint32), b, x, y, p, nthreads)
    return y


def krylov(A, p, b, dr, dc, x_in, t_begin, t_end, want_x=False, nthreads=0):
    P = b * (b + 1) // 2
    seq = np.zeros((max(t_end - t_begin, 1), P), np.uint32)
    x_out = np.zeros((A.n, b), np.uint32) if want_x else None
    rc = lib().orc_krylov(A.m, A.n, _c(A.row_ptr, np.uint64), _c(A.col, np.uint32), _c(A.val, np.uint32), p, b,
                          _c(dr, np.uint32), _c(dc, np.uint32), _c(x_in, np.uint32), t_begin, t_end,
                          seq, x_out.ctypes.data if want_x else None, nthreads)
    if rc != 0:
        raise RuntimeError(f"orc_krylov failed: {rc}")
    seq = seq[: t_end - t_begin]
    return (seq, x_out) if want_x else seq


def dense_rank(mat, p, nthreads=0):
    a = np.array(mat, dtype=np.uint32, copy=True, order="C")
    return int(lib().orc_dense_rank(a.shape[0], a.shape[1], a, p, nthreads))


def ntt(a, invert=False):
    a = np.array(a, dtype=np.uint32, copy=True)
    assert lib().orc_ntt(a, a.size, int(invert)) == 0
    return a


def ntt_generic(a, p, g, invert=False):
    a = np.array(a, dtype=np.uint64, copy=True)
    assert lib().orc_ntt_generic(a, a.size, int(invert), p, g) == 0
    return a


def lsp(a, p):
    a = _c(a, np.uint32)
    m = a.shape[0]
    L = np.zeros((m, m), np.uint32)
    S = np.zeros((m, m), np.uint32)
    perm = np.zeros(m, np.uint32)
    rank = lib().orc_lsp(m, a, p, L, S, perm)
    return L, S, perm, rank


def matrix_inverse(a, p):
    a = _c(a, np.uint32)
    out = np.zeros_like(a)
    rc = lib().orc_matrix_inverse(a.shape[0], a, p, out)
    return None if rc != 0 else out


def modular_rank(a, p):
    a = _c(a, np.uint32)
    return int(lib().orc_modular_rank(a.shape[0], a.shape[1], a, p))


def poly_mat_mul(A, B, lo, hi, p, nb_lim=None, use_ntt=True):
    A = _c(A, np.uint32)
    B = _c(B, np.uint32)
    m, n, na = A.shape
    n2, k, nb = B.shape
    assert n == n2
    if nb_lim is not None and nb_lim < nb:
        B = np.ascontiguousarray(B[:, :, :nb_lim])
        nb = nb_lim
    out = np.zeros((m, k, hi - lo), np.uint32)
    lib().orc_set_use_ntt(int(use_ntt))
    lib().orc_poly_mat_mul(m, n, k, A, na, B, nb, lo, hi, p, out)
    lib().orc_set_use_ntt(1)
    return out


def m_basis(G0, delta, p):
    G0 = _c(G0, np.uint32)
    m, n = G0.shape
    d = np.array(delta, dtype=np.int64, copy=True)
    M = np.zeros((m, m, 2), np.uint32)
    lib().orc_m_basis(m, n, G0, d, p, M)
    return M, d


def pm_basis(seq, b, d, p):
    seq = _c(seq, np.uint32)
    L = seq.shape[0]
    basis = np.zeros((2 * b, 2 * b, d + 1), np.uint32)
    delta = np.zeros(2 * b, np.int64)
    rc = lib().orc_pm_basis(seq, b, L, d, p, basis, delta)
    if rc != 0:
        raise ValueError(f"orc_pm_basis rejected d={d}, L={L}")
    return basis, delta


def rank_from_basis(basis, delta, b, p):
    basis = _c(basis, np.uint32)
    return int(lib().orc_rank_from_basis(basis, _c(delta, np.int64), b, basis.shape[2], p))


# ---------------------------------------------------------------- reference calls
def ref_ntt_p61(a, invert=False):
    a = np.array(a, dtype=np.uint64, copy=True)
    ref().ref_ntt_u64(a, a.size, int(invert))
    return a


def ref_poly_mat_mul(A, B, b_start, b_end, p):
    A = _c(A, np.uint32)
    B = _c(B, np.uint32)
    m, n, na = A.shape
    _, k, nb = B.shape
    nbb = min(nb, b_end - b_start)
    out = np.zeros((m, k, na + nbb - 1), np.uint32)
    ref().ref_poly_mat_mul_u64(m, n, k, A, na, B, nb, b_start, b_end, p, out)
    return out


def ref_lsp(a, p):
    a = _c(a, np.uint32)
    m = a.shape[0]
    L = np.zeros((m, m), np.uint32)
    S = np.zeros((m, m), np.uint32)
    P = np.zeros((m, m), np.uint32)
    ref().ref_lsp_u64(m, a, p, L, S, P)
    return L, S, P


def ref_inverse(a, p):
    a = _c(a, np.uint32)
    out = np.zeros_like(a)
    return None if ref().ref_inverse_u64(a.shape[0], a, p, out) != 0 else out


def ref_modular_rank(a, p):
    a = _c(a, np.uint32)
    return int(ref().ref_modular_rank_u64(a.shape[0], a.shape[1], a, p))


def ref_m_basis(G0, delta, p):
    G0 = _c(G0, np.uint32)
    m, n = G0.shape
    d = np.array(delta, dtype=np.int64, copy=True)
    M = np.zeros((m, m, 2), np.uint32)
    ref().ref_m_basis_u64(m, n, G0, d, p, M)
    return M, d


def ref_pm_basis(seq, b, d, p):
    seq = _c(seq, np.uint32)
    L = seq.shape[0]
    dd = 1
    while dd < d:
        dd *= 2
    cap = 4 * b * b * (dd + 1)
    basis = np.zeros(cap, np.uint32)
    delta = np.zeros(2 * b, np.int64)
    ln = ref().ref_pm_basis_u64(seq, b, L, d, p, basis, cap, delta)
    assert ln > 0
    return basis[: 4 * b * b * ln].reshape(2 * b, 2 * b, ln), delta


def ref_rank(basis, delta, b, p):
    basis = _c(basis, np.uint32)
    return int(ref().ref_analyze_pm_basis_u64(basis, _c(delta, np.int64), b, basis.shape[2], p))


def ref_load_sms(path, p):
    m, n, nnz = C.c_int64(), C.c_int64(), C.c_int64()
    R = ref()
    R.ref_load_sms(path.encode(), p, C.byref(m), C.byref(n), C.byref(nnz), None, None, None)
    rp = np.zeros(m.value + 1, np.int64)
    col = np.zeros(nnz.value, np.int32)
    val = np.zeros(nnz.value, np.uint32)
    R.ref_load_sms(path.encode(), p, C.byref(m), C.byref(n), C.byref(nnz),
                   rp.ctypes.data, col.ctypes.data, val.ctypes.data)
    return m.value, n.value, rp, col, val


def ref_transpose(m, n, rp, col, val):
    nnz = int(rp[m])
    rpt = np.zeros(n + 1, np.int64)
    colt = np.zeros(nnz, np.int32)
    valt = np.zeros(nnz, np.uint32)
    ref().ref_transpose(m, n, _c(rp, np.int64), _c(col, np.int32), _c(val, np.uint32), rpt, colt, valt)
    return rpt, colt, valt


def ref_rescale(m, n, rp, col, val, dr, dc, p):
    v = np.array(val, dtype=np.uint32, copy=True)
    drc = _c(dr, np.uint32) if dr is not None else None
    dcc = _c(dc, np.uint32) if dc is not None else None
    ref().ref_rescale(m, n, _c(rp, np.int64), _c(col, np.int32), v,
                      drc.ctypes.data if drc is not None else None,
                      dcc.ctypes.data if dcc is not None else None, p)
    return v


def ref_save_wdm(path, m, n, p, dr, dc, V, curv, seq):
    """The reference's save_wdm_file_sym (cuda_test/wdmfiles.h:92-160); V, curv n x b, seq [len][P]."""
    V = _c(V, np.uint32)
    b = V.shape[1]
    seq = _c(seq, np.uint32).reshape(-1, b * (b + 1) // 2)
    ref().ref_save_wdm(os.fsencode(path), m, n, p, b, _c(dr, np.uint32), _c(dc, np.uint32), V, _c(curv, np.uint32),
                       seq, seq.shape[0])
