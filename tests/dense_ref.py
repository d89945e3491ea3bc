"""Independent dense arithmetic for tests: exact mod-p products with numpy
uint64 by splitting one operand into 16-bit limbs (no call into the oracle)."""
import numpy as np


def matmul_mod(A, B, p):
    """(A @ B) mod p for entries in [0, p), p < 2^31, inner dimension < 2^9."""
    A = np.asarray(A, dtype=np.uint64)
    B = np.asarray(B, dtype=np.uint64)
    assert A.shape[1] < 512
    lo = B & np.uint64(0xFFFF)
    hi = B >> np.uint64(16)
    P = np.uint64(p)
    r_lo = (A @ lo) % P
    r_hi = (A @ hi) % P
    return ((r_hi * np.uint64(1 << 16) + r_lo) % P).astype(np.uint64)


def krylov_dense(Ad, dr, dc, V, p, L):
    """S_t = V^T B^t V with B = D_c A^T D_r A D_c, upper triangles in row-major
    pair order, t = 0..L-1 (the sequence of src/main.rs:88-216)."""
    m, n = Ad.shape
    b = V.shape[1]
    P = np.uint64(p)
    a = (Ad.astype(np.uint64) * dr.astype(np.uint64)[:, None]) % P
    a = (a * dc.astype(np.uint64)[None, :]) % P            # D_r A D_c
    at = (Ad.T.astype(np.uint64) * dc.astype(np.uint64)[:, None]) % P  # D_c A^T
    iu = np.triu_indices(b)
    xs = [V.astype(np.uint64)]
    need = (L + 1) // 2 + 1
    for _ in range(need):
        y = matmul_mod(a, xs[-1], p)
        xs.append(matmul_mod(at, y, p))
    out = []
    for t in range(L):
        k1, k2 = t // 2, t - t // 2
        G = matmul_mod(xs[k1].T.copy(), xs[k2], p)
        out.append(G[iu])
    return np.array(out, dtype=np.uint32), xs
