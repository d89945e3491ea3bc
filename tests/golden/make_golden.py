"""Writes the golden vectors in tests/golden/*.npz from the reference's own CPU
C++ (cuda_test/*.h), compiled by oracle/Makefile into oracle/_ref/libref.so.

Run in a container that holds the reference checkout:
    make -C oracle && python tests/golden/make_golden.py

Every fixture is data: inputs and the reference's outputs for them.
  ntt_p61.npz        ntt<uint64_t> (cuda_test/ntt.h:152-193), mod 2^61-2^26+1, root 3
  poly_mat_mul.npz   poly_mat_mul_fft_red<uint64_t> (cuda_test/poly_mat_mul.h:227-251) at small p
  linalg.npz         lsp_decomposition / matrix_inverse / modular_rank (cuda_test/modular_linalg.h)
                     on the known-answer matrices of src/modular_linalg.rs:437-507 and random ones
  m_basis.npz        M_Basis<uint64_t> (cuda_test/sigma_basis.h:91-171), deltas with ties
  pm_basis.npz       PM_Basis<uint64_t,uint64_t> + analyze_pm_basis (sigma_basis.h:214-360)
                     on Krylov sequences of small matrices at p in {7, 97, 3323, 65521}
  matrix_txt.npz     cuda_test/matrix.txt through the reference SMS loader, transpose and
                     rescaling (cuda_test/matrices.h), and its dense rank (modular_rank)
  wdm_ref.npz        .wdm files written by the reference's save_wdm_file_sym
                     (cuda_test/wdmfiles.h:92-160): the inputs and the file's bytes
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)

import pyoracle as O  # noqa: E402
from rust_rank_amd import synth  # noqa: E402

REF_MATRIX = "/root/reference/cuda_test/matrix.txt"
P61 = 2305843009146585089
P = 2013265921


def save(name, **arrays):
    np.savez_compressed(os.path.join(HERE, name), **arrays)
    print("wrote", name, sum(a.nbytes for a in arrays.values()), "bytes raw")


def main():
    rng = np.random.default_rng(20251203)

    # ---- NTT mod P61: deterministic i % p data (cudantt.h:436-494 style) and random
    out = {}
    for i, n in enumerate([2, 16, 256, 4096]):
        a = (np.arange(n, dtype=np.uint64) * np.uint64(2654435761) + np.uint64(i)) % np.uint64(P61)
        if i == 3:
            a = rng.integers(0, P61, n, dtype=np.uint64)
        fwd = O.ref_ntt_p61(a)
        inv = O.ref_ntt_p61(fwd, invert=True)
        out[f"in{i}"], out[f"fwd{i}"], out[f"inv{i}"] = a, fwd, inv
    save("ntt_p61.npz", **out)

    # ---- polynomial-matrix products at small p
    out = {}
    cases = [(7, 2, 2, 2, 3, 5, 0, 5), (97, 4, 4, 2, 7, 13, 2, 11), (3323, 4, 8, 4, 9, 17, 0, 17),
             (65521, 3, 5, 2, 33, 40, 5, 30)]
    for i, (p, m, n, k, na, nb, b0, b1) in enumerate(cases):
        A = rng.integers(0, p, (m, n, na)).astype(np.uint32)
        B = rng.integers(0, p, (n, k, nb)).astype(np.uint32)
        Cr = O.ref_poly_mat_mul(A, B, b0, b1, p)
        out.update({f"p{i}": np.array([p, b0, b1], np.int64), f"A{i}": A, f"B{i}": B, f"C{i}": Cr})
    save("poly_mat_mul.npz", **out)

    # ---- dense linear algebra
    out = {}
    mats = [
        (101, np.array([[2, 4, 6], [1, 3, 5], [0, 0, 1]])),
        (97, np.array([[10, 20, 30, 40, 50], [5, 10, 15, 20, 25], [0, 1, 0, 1, 0], [1, 0, 1, 0, 1],
                       [99, 99, 99, 99, 99]])),
        (97, np.array([[2, 3, 1], [1, 1, 1], [3, 5, 2]])),
        (97, np.eye(4, dtype=np.int64)),
    ]
    for p, m in [(97, 8), (3323, 12), (P, 16), (7, 10)]:
        a = rng.integers(0, p, (m, m))
        a[2] = (3 * a[0] + 5 * a[1]) % p
        mats.append((p, a))
    for i, (p, a) in enumerate(mats):
        a = (np.asarray(a) % p).astype(np.uint32)
        L, S, Pm = O.ref_lsp(a, p)
        inv = O.ref_inverse(a, p)
        rk = O.ref_modular_rank(a, p)
        out.update({f"p{i}": np.array([p]), f"a{i}": a, f"L{i}": L, f"S{i}": S, f"P{i}": Pm,
                    f"rank{i}": np.array([rk]), f"inv{i}": inv if inv is not None else np.zeros(0, np.uint32)})
    # a unit lower-triangular L from an LSP, inverted (the use in M_Basis)
    save("linalg.npz", **out)

    # ---- M_Basis leaves, with repeated deltas so the stable sort matters
    out = {}
    i = 0
    for p in (7, 97, 3323, P):
        for b in (1, 2, 4, 8):
            G0 = rng.integers(0, p, (2 * b, b)).astype(np.uint32)
            if b > 1:
                G0[1] = 0
                G0[2 * b - 1] = G0[0]
            delta = rng.integers(-3, 1, 2 * b).astype(np.int64)
            M, d2 = O.ref_m_basis(G0, delta, p)
            out.update({f"p{i}": np.array([p]), f"G{i}": G0, f"din{i}": delta, f"M{i}": M, f"dout{i}": d2})
            i += 1
    save("m_basis.npz", **out)

    # ---- PM_Basis + rank on Krylov sequences of small sparse matrices
    out = {}
    cases = [(7, 2, 20, 16, 4), (97, 2, 40, 30, 3), (3323, 2, 60, 50, 3), (65521, 4, 120, 100, 5),
             (97, 4, 90, 80, 6), (3323, 8, 300, 260, 10)]
    for i, (p, b, m, n, dep) in enumerate(cases):
        A = synth.uniform_sparse(11 + i, 1, m, n, 2, 5, p, small=False)
        A = synth.plant_dependent_columns(synth.Rng(11 + i, 2), A, dep, p, small=False)
        dr, dc, V = synth.preconditioners(11 + i, m, n, b, p)
        L = 2 * n // b + 8
        d = 1
        while d * 2 < L:
            d *= 2
        seq = O.krylov(A, p, b, dr, dc, V, 0, L)
        basis, delta = O.ref_pm_basis(seq, b, d, p)
        rk = O.ref_rank(basis, delta, b, p)
        out.update({f"meta{i}": np.array([p, b, d, L], np.int64), f"seq{i}": seq, f"basis{i}": basis,
                    f"delta{i}": delta, f"rank{i}": np.array([rk], np.int64)})
    # test_cuPMBasis shape (cusigma_basis.h:325-369): p = 7, b = 2, d = 4 on values in [0, 5)
    seq = rng.integers(0, 5, (9, 3)).astype(np.uint32)
    basis, delta = O.ref_pm_basis(seq, 2, 4, 7)
    i = len(cases)
    out.update({f"meta{i}": np.array([7, 2, 4, 9], np.int64), f"seq{i}": seq, f"basis{i}": basis,
                f"delta{i}": delta, f"rank{i}": np.array([O.ref_rank(basis, delta, 2, 7)], np.int64)})
    save("pm_basis.npz", **out)

    # ---- cuda_test/matrix.txt through the reference loader, transpose, rescaling, rank
    if os.path.exists(REF_MATRIX):
        m, n, rp, col, val = O.ref_load_sms(REF_MATRIX, P)
        rpt, colt, valt = O.ref_transpose(m, n, rp, col, val)
        dr, dc, _ = synth.preconditioners(77, m, n, 1, P)
        a_scaled = O.ref_rescale(m, n, rp, col, val, dr, dc, P)        # D_r A D_c
        at_scaled = O.ref_rescale(n, m, rpt, colt, valt, dc, None, P)  # D_c A^T
        dense = np.zeros((m, n), np.uint64)
        rows = np.repeat(np.arange(m), np.diff(rp))
        np.add.at(dense, (rows, col), val.astype(np.uint64))
        dense = (dense % np.uint64(P)).astype(np.uint32)
        rk = O.ref_modular_rank(dense, P)
        save("matrix_txt.npz", shape=np.array([m, n]), row_ptr=rp, col=col, val=val, row_ptr_t=rpt, col_t=colt,
             val_t=valt, dr=dr, dc=dc, a_scaled=a_scaled, at_scaled=at_scaled, rank=np.array([rk]))


def wdm_fixture():
    """Inputs and the bytes of the .wdm file the reference's writer makes of them."""
    import tempfile
    rng = np.random.default_rng(2024)
    out = {}
    for i, (p, m, n, b, ln) in enumerate([(3323, 7, 5, 3, 4), (P, 6, 4, 2, 5), (97, 3, 9, 1, 1)]):
        dr = rng.integers(1, p, m).astype(np.uint32)
        dc = rng.integers(1, p, n).astype(np.uint32)
        V = rng.integers(0, p, (n, b)).astype(np.uint32)
        curv = rng.integers(0, p, (n, b)).astype(np.uint32)
        seq = rng.integers(0, p, (ln, b * (b + 1) // 2)).astype(np.uint32)
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "ref.wdm")
            O.ref_save_wdm(path, m, n, p, dr, dc, V, curv, seq)
            data = np.frombuffer(open(path, "rb").read(), dtype=np.uint8)
        out.update({f"meta{i}": np.array([p, m, n, b, ln], np.int64), f"dr{i}": dr, f"dc{i}": dc, f"V{i}": V,
                    f"curv{i}": curv, f"seq{i}": seq, f"file{i}": data})
    save("wdm_ref.npz", **out)


if __name__ == "__main__":
    if sys.argv[1:] == ["wdm"]:
        wdm_fixture()
    else:
        main()
        wdm_fixture()
