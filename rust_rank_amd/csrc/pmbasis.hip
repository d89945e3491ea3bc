// pmbasis.hip -- PM-basis (order-d approximant basis of G = [S(x); I_b]) and the
// rank estimate, on the device.
//
// Replaces PM_Basis / _PM_Basis / M_Basis / analyze_pm_basis
// (src/sigma_basis.rs:9-54, 190-219, 348-416, 71-152) with the rank formula of
// src/blockbmtest.rs:574-600; the CUDA twin cuda_test/cusigma_basis.h:131-320
// copies every leaf to the host, here nothing leaves the device until the
// basis is done.
//
// The order d is arbitrary (the reference needs a power of two,
// src/sigma_basis.rs:22-29): it splits as d1 = d - d/2, d2 = d/2, which is the
// reference's split for powers of two.  The basis is the product of the
// order-1 leaves M_{d-1} ... M_0, each computed from the constant term of the
// exact residual, so any split gives the same bits.
//
// Leaf (M_Basis): one workgroup.  Rows of the constant term are sorted by
// delta (stable, descending), then eliminated LSP-style (pivot = first nonzero
// position >= rank in the swapped column order, src/modular_linalg.rs:188-239)
// fraction-free, so no pivot needs an inverse inside the loop: every row
// carries its multiplier sigma_i and its coefficients on earlier pivot rows.
// One parallel inversion of the sigmas at the end gives Linv, and
// M = D1 Linv pi + x Dx Linv pi (src/sigma_basis.rs:384-396).
//
// Products: schoolbook kernels for short operands, otherwise forward NTTs of
// both operands, pointwise 2b x 2b products and inverse NTTs that write only
// the needed coefficient window (the middle product of the first half needs a
// cyclic transform of d+1 points, not 1.5d).
#include "common.hpp"
#include "ntt.hpp"
#include <vector>
#include <algorithm>
#include <cstring>

namespace rr {

// ---------------------------------------------------------------------------
// leaf kernel
// G0(i, j) = G[(i * n + j) * gstride + goff]; M [m][m][2] (stride 2); delta [m] in/out.
__global__ __launch_bounds__(256) void k_mbasis(const uint32_t *G, uint64_t gstride, uint64_t goff, uint32_t m,
                                                uint32_t n, int64_t *delta, uint32_t *M, uint32_t p, double pd,
                                                double inv) {
    extern __shared__ uint32_t sm[];
    uint32_t *R = sm;                  // [m][n] reduced rows
    uint32_t *U = R + m * n;           // [m][n] coefficients on pivot rows
    uint32_t *sig = U + m * n;         // [m]
    uint32_t *fcol = sig + m;          // [m] snapshot of the pivot column
    uint32_t *perm = fcol + m;         // [m]
    uint32_t *colord = perm + m;       // [m] column at each position
    uint32_t *pivrow = colord + m;     // [n]
    uint32_t *ctl = pivrow + n;        // [4]: pivot position, pivot column, rank
    int64_t *dl = (int64_t *)(ctl + 4 + ((m + n) & 1));  // [m] (8-byte aligned)
    const uint32_t tid = threadIdx.x, nt = blockDim.x;

    for (uint32_t i = tid; i < m; i += nt) dl[i] = delta[i];
    __syncthreads();
    // stable descending sort of delta: rank_i = #{j: dl_j > dl_i} + #{j < i: dl_j == dl_i}
    for (uint32_t i = tid; i < m; i += nt) {
        int64_t di = dl[i];
        uint32_t r = 0;
        for (uint32_t j = 0; j < m; j++) r += (dl[j] > di) || (dl[j] == di && j < i);
        perm[r] = i;
    }
    for (uint32_t i = tid; i < m; i += nt) {
        colord[i] = i;
        sig[i] = 1;
    }
    __syncthreads();
    for (uint32_t idx = tid; idx < m * n; idx += nt) {
        uint32_t i = idx / n, j = idx % n;
        R[idx] = G[((size_t)perm[i] * n + j) * gstride + goff] % p;
        U[idx] = 0;
    }
    if (tid == 0) ctl[2] = 0;
    __syncthreads();

    for (uint32_t k = 0; k < m; k++) {
        uint32_t rank = ctl[2];
        if (rank >= n) break;  // every later row is zero: rank <= n
        if (tid == 0) {
            uint32_t pos = m;
            for (uint32_t q = rank; q < m; q++) {
                uint32_t c = colord[q];
                if (c < n && R[k * n + c] != 0) { pos = q; break; }
            }
            ctl[0] = pos;
            if (pos < m) {
                uint32_t c = colord[pos];
                colord[pos] = colord[rank];
                colord[rank] = c;
                ctl[1] = c;
                pivrow[rank] = k;
            }
        }
        __syncthreads();
        const uint32_t pos = ctl[0], c = ctl[1];
        __syncthreads();  // ctl is rewritten by thread 0 in the next iteration
        if (pos >= m) continue;  // zero row (uniform branch)
        for (uint32_t i = k + 1 + tid; i < m; i += nt) fcol[i] = R[i * n + c];
        __syncthreads();
        const uint32_t pk = R[k * n + c], sk = sig[k];
        const uint32_t W = n + rank + 1;  // R columns, U columns < rank, the new U column
        for (uint32_t idx = tid; idx < (m - k - 1) * W; idx += nt) {
            uint32_t i = k + 1 + idx / W, col = idx % W;
            uint32_t f = fcol[i];
            if (f == 0) continue;
            if (col < n) {
                R[i * n + col] = submod(mulmod(pk, R[i * n + col], p, inv), mulmod(f, R[k * n + col], p, inv), p);
            } else if (col < n + rank) {
                uint32_t q = col - n;
                U[i * n + q] = submod(mulmod(pk, U[i * n + q], p, inv), mulmod(f, U[k * n + q], p, inv), p);
            } else {
                U[i * n + rank] = submod(0u, mulmod(f, sk, p, inv), p);
                sig[i] = mulmod(pk, sig[i], p, inv);
            }
        }
        if (tid == 0) ctl[2] = rank + 1;
        __syncthreads();
    }
    const uint32_t rank = ctl[2];
    // sigma_i^-1 in parallel (Fermat), then Linv row i = coefficients / sigma_i
    for (uint32_t i = tid; i < m; i += nt) {
        uint32_t s = sig[i], r = 1, e = p - 2;
        while (e) {
            if (e & 1) r = mulmod(r, s, p, inv);
            s = mulmod(s, s, p, inv);
            e >>= 1;
        }
        sig[i] = r;
    }
    __syncthreads();
    // row i is nonzero iff it was a pivot row
    for (uint32_t idx = tid; idx < m * m; idx += nt) {
        uint32_t i = idx / m, j = idx % m;
        M[(size_t)idx * 2] = 0;
        M[(size_t)idx * 2 + 1] = 0;
        (void)j;
    }
    __syncthreads();
    for (uint32_t i = tid; i < m; i += nt) {
        bool nz = false;
        for (uint32_t q = 0; q < rank; q++) nz |= pivrow[q] == i;
        uint32_t slot = nz ? 1 : 0;
        // T = Linv pi: T[i][perm[j]] = Linv[i][j]
        M[((size_t)i * m + perm[i]) * 2 + slot] = 1;
        for (uint32_t q = 0; q < rank; q++) {
            uint32_t k = pivrow[q];
            if (k >= i) continue;
            uint32_t v = mulmod(U[i * n + q], sig[i], p, inv);
            M[((size_t)i * m + perm[k]) * 2 + slot] = v;
        }
        delta[i] = dl[perm[i]] - (nz ? 1 : 0);
    }
}

// ---------------------------------------------------------------------------
struct Pool {  // bump allocator over one device buffer
    DevBuf buf;
    size_t top = 0, cap = 0;
    rr_status init(size_t words) {
        cap = words;
        top = 0;
        return buf.alloc(words * 4);
    }
    uint32_t *get(size_t words) {
        words = (words + 63) & ~size_t(63);
        if (top + words > cap) return nullptr;
        uint32_t *p = buf.as<uint32_t>() + top;
        top += words;
        return p;
    }
};

struct PmCtx {
    uint32_t p, b, m;  // m = 2b
    hipStream_t st;
    NttTables tables;
    bool use_ntt;
    Pool pool;
    DevBuf fa, fb, fc;  // spectra workspace
    size_t fcap = 0;
    DevBuf delta;
    size_t lds_bytes;
};

// C [rows][cols][hi-lo] (contiguous) = window [lo, hi) of A [rows][inner][la] * B [inner][cols][ob .. ob+lb)
static rr_status pmat_mul(PmCtx &c, const uint32_t *A, uint32_t la, const uint32_t *B, uint64_t sb, uint32_t ob,
                          uint32_t lb, uint32_t *C, uint32_t rows, uint32_t inner, uint32_t cols, uint32_t lo,
                          uint32_t hi) {
    uint64_t school_cost = (uint64_t)std::min(la, hi) * std::min(lb, hi);
    if (!c.use_ntt || school_cost <= 256) {
        return pmul_school(A, la, la, B, sb, ob, lb, C, hi - lo, rows, inner, cols, lo, hi, c.p, c.st);
    }
    uint32_t need = std::max<uint32_t>(hi, la + lb - 1 - lo);
    uint32_t logn = 1;
    while ((1u << logn) < need) logn++;
    RR_CHECK(logn <= NTT_MAX_LOGN, RR_EINVAL, "product needs a 2^%u-point transform (max 2^%u)", logn,
             NTT_MAX_LOGN);
    const uint32_t N = 1u << logn;
    size_t na = (size_t)rows * inner * N, nb = (size_t)inner * cols * N, nc = (size_t)rows * cols * N;
    RR_CHECK(na <= c.fcap && nb <= c.fcap && nc <= c.fcap, RR_ENOMEM, "spectrum workspace too small");
    rr_status s;
    if ((s = ntt_forward(c.tables, A, la, 0, la, c.fa.as<uint32_t>(), rows * inner, logn, c.st)) != RR_OK) return s;
    if ((s = ntt_forward(c.tables, B, sb, ob, lb, c.fb.as<uint32_t>(), inner * cols, logn, c.st)) != RR_OK) return s;
    uint32_t ninv = h_invmod(N, NTT_PRIME);
    if ((s = pointwise(c.fa.as<uint32_t>(), c.fb.as<uint32_t>(), c.fc.as<uint32_t>(), rows, inner, cols, N, ninv,
                       c.st)) != RR_OK)
        return s;
    return ntt_inverse(c.tables, c.fc.as<uint32_t>(), rows * cols, logn, lo, hi - lo, C, hi - lo, c.st);
}

// M [m][m][d+1] = order-d basis of G (m x b polys, stride gs, coefficients 0..d valid).
static rr_status pm_rec(PmCtx &c, const uint32_t *G, uint64_t gs, uint64_t d, uint32_t *M) {
    const uint32_t m = c.m, b = c.b;
    if (d == 1) {
        hipLaunchKernelGGL(k_mbasis, dim3(1), dim3(256), c.lds_bytes, c.st, G, gs, (uint64_t)0, m, b,
                           c.delta.as<int64_t>(), M, c.p, (double)c.p, 1.0 / (double)c.p);
        RR_HIP(hipGetLastError());
        return RR_OK;
    }
    const uint64_t d1 = d - d / 2, d2 = d / 2;
    size_t mark = c.pool.top;
    uint32_t *M1 = c.pool.get((size_t)m * m * (d1 + 1));
    uint32_t *GG = c.pool.get((size_t)m * b * (d2 + 1));
    uint32_t *M2 = c.pool.get((size_t)m * m * (d2 + 1));
    RR_CHECK(M1 && GG && M2, RR_ENOMEM, "PM-basis workspace exhausted");
    rr_status s;
    if ((s = pm_rec(c, G, gs, d1, M1)) != RR_OK) return s;
    // GG = coefficients d1..d of M1 * G[0..d]   (src/sigma_basis.rs:205-209)
    if ((s = pmat_mul(c, M1, (uint32_t)(d1 + 1), G, gs, 0, (uint32_t)(d + 1), GG, m, m, b, (uint32_t)d1,
                      (uint32_t)(d + 1))) != RR_OK)
        return s;
    if ((s = pm_rec(c, GG, d2 + 1, d2, M2)) != RR_OK) return s;
    // M = M2 * M1   (src/sigma_basis.rs:213)
    if ((s = pmat_mul(c, M2, (uint32_t)(d2 + 1), M1, d1 + 1, 0, (uint32_t)(d1 + 1), M, m, m, m, 0,
                      (uint32_t)(d + 1))) != RR_OK)
        return s;
    c.pool.top = mark;
    return RR_OK;
}

static uint64_t rec_words(uint64_t d, uint32_t m, uint32_t b) {
    if (d <= 1) return 0;
    uint64_t d1 = d - d / 2, d2 = d / 2;
    uint64_t here = ((uint64_t)m * m * (d1 + 1) + 63) + ((uint64_t)m * b * (d2 + 1) + 63) +
                    ((uint64_t)m * m * (d2 + 1) + 63);
    return here + std::max(rec_words(d1, m, b), rec_words(d2, m, b));
}

// Device PM-basis from a device-resident G [m][b][d+1]; writes M [m][m][d+1] (device) and delta (host).
rr_status pm_basis_device(uint32_t p, uint32_t b, uint64_t d, const uint32_t *G, uint32_t *M, int64_t *delta_host,
                          hipStream_t st) {
    PmCtx c;
    c.p = p;
    c.b = b;
    c.m = 2 * b;
    c.st = st;
    c.use_ntt = (p == NTT_PRIME);
    const uint32_t m = c.m;
    c.lds_bytes = (size_t)(2 * m * b + 4 * m + b + 6) * 4 + (size_t)m * 8 + 16;
    if (c.lds_bytes > 65536)
        RR_HIP(hipFuncSetAttribute((const void *)k_mbasis, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)c.lds_bytes));
    rr_status s;
    if ((s = c.delta.alloc(m * 8)) != RR_OK) return s;
    RR_HIP(hipMemsetAsync(c.delta.ptr, 0, m * 8, st));
    if ((s = c.pool.init(rec_words(d, m, b) + 1024)) != RR_OK) return s;
    if (c.use_ntt) {
        uint32_t logn = 1;
        while ((1ull << logn) < d + 1) logn++;
        if (logn > NTT_MAX_LOGN) logn = NTT_MAX_LOGN;
        if ((s = c.tables.init(std::max<uint32_t>(logn, 2), st)) != RR_OK) return s;
        c.fcap = (size_t)m * m << logn;
        if ((s = c.fa.alloc(c.fcap * 4)) != RR_OK) return s;
        if ((s = c.fb.alloc(c.fcap * 4)) != RR_OK) return s;
        if ((s = c.fc.alloc(c.fcap * 4)) != RR_OK) return s;
    }
    if ((s = pm_rec(c, G, d + 1, d, M)) != RR_OK) return s;
    RR_HIP(hipMemcpyAsync(delta_host, c.delta.ptr, m * 8, hipMemcpyDeviceToHost, st));
    RR_HIP(hipStreamSynchronize(st));
    return RR_OK;
}

// src/modular_linalg.rs:132-183 on a b x b host matrix
static uint32_t host_rank(std::vector<uint32_t> a, uint32_t rows, uint32_t cols, uint32_t p) {
    uint32_t rank = 0;
    for (uint32_t c = 0; c < cols && rank < rows; c++) {
        uint32_t piv = rank;
        while (piv < rows && a[(size_t)piv * cols + c] == 0) piv++;
        if (piv == rows) continue;
        for (uint32_t j = 0; j < cols; j++) std::swap(a[(size_t)rank * cols + j], a[(size_t)piv * cols + j]);
        uint32_t iv = h_invmod(a[(size_t)rank * cols + c], p);
        for (uint32_t j = c; j < cols; j++) a[(size_t)rank * cols + j] = h_mulmod(a[(size_t)rank * cols + j], iv, p);
        for (uint32_t r = rank + 1; r < rows; r++) {
            uint32_t f = a[(size_t)r * cols + c];
            if (!f) continue;
            for (uint32_t j = c; j < cols; j++) {
                uint32_t sub = h_mulmod(f, a[(size_t)rank * cols + j], p);
                uint32_t &x = a[(size_t)r * cols + j];
                x = x >= sub ? x - sub : x + p - sub;
            }
        }
        rank++;
    }
    return rank;
}

// analyze_pm_basis + analyze_min_generators (src/sigma_basis.rs:115-145, src/blockbmtest.rs:598),
// with the reference's unsigned wrap when the reversed generator has fewer than 3 coefficients.
int64_t rank_from_basis_host(const uint32_t *basis, const int64_t *delta, uint32_t b, uint64_t len, uint32_t p,
                             rr_status *st) {
    const uint32_t n2 = 2 * b;
    uint64_t maxdel = 0;
    for (uint32_t i = 0; i < b; i++) {
        if (delta[i] > 0 || (uint64_t)(-delta[i]) >= len) {
            *st = RR_EINVAL;
            set_error("delta[%u]=%lld outside [-(len-1), 0]", i, (long long)delta[i]);
            return 0;
        }
        maxdel = std::max<uint64_t>(maxdel, (uint64_t)(-delta[i]));
    }
    std::vector<uint32_t> F((maxdel + 1) * b * b, 0);
    uint64_t maxdeg = 0;
    for (uint32_t i = 0; i < b; i++) {
        uint64_t del = (uint64_t)(-delta[i]);
        for (uint32_t j = 0; j < b; j++)
            for (uint64_t k = 0; k <= del; k++) {
                uint32_t v = basis[((size_t)i * n2 + j) * len + (del - k)];
                F[k * b * b + i * b + j] = v;
                if (v && k > maxdeg) maxdeg = k;
            }
    }
    uint64_t K = maxdeg + 1;
    std::vector<uint32_t> last(F.begin() + (K - 1) * b * b, F.begin() + K * b * b), first(F.begin(), F.begin() + b * b);
    uint64_t r = (uint64_t)b * (K - 3) + host_rank(last, b, b, p) + host_rank(first, b, b, p);
    *st = RR_OK;
    return (int64_t)r;
}

// G [2b][b][d+1] from the t-major upper-triangle sequence [L][P] (src/sigma_basis.rs:33-46)
__global__ void k_build_G(const uint32_t *seq, uint32_t b, uint64_t len, uint32_t *G) {
    const uint32_t P = b * (b + 1) / 2;
    const uint64_t total = (uint64_t)2 * b * b * len;
    for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t k = idx % len, ij = idx / len;
        uint32_t j = (uint32_t)(ij % b), i = (uint32_t)(ij / b);
        uint32_t v;
        if (i < b) {
            uint32_t lo = min(i, j), hi = max(i, j);
            uint32_t ii = lo * b - lo * (lo - 1) / 2 + (hi - lo);
            v = seq[k * P + ii];
        } else {
            v = (k == 0 && i - b == j) ? 1u : 0u;
        }
        G[idx] = v;
    }
}

rr_status build_G_device(const uint32_t *seq_dev, uint32_t b, uint64_t len, uint32_t *G, hipStream_t st) {
    uint64_t total = (uint64_t)2 * b * b * len;
    uint32_t blocks = (uint32_t)std::min<uint64_t>((total + 255) / 256, 16384);
    hipLaunchKernelGGL(k_build_G, dim3(std::max(1u, blocks)), dim3(256), 0, st, seq_dev, b, len, G);
    RR_HIP(hipGetLastError());
    return RR_OK;
}

}  // namespace rr

using namespace rr;

extern "C" {

rr_status rr_pm_basis(const uint32_t *seq, uint32_t b, uint64_t L, uint64_t d, uint32_t p, uint32_t *basis_out,
                      int64_t *delta_out) {
    clear_error();
    RR_CHECK(seq && basis_out && delta_out, RR_EINVAL, "null argument");
    RR_CHECK(b >= 1 && b <= 64, RR_EINVAL, "b=%u must be in 1..64", b);
    RR_CHECK(p > 2 && p < (1u << 31) && h_is_prime(p), RR_EINVAL, "p=%u must be a prime in (2, 2^31)", p);
    RR_CHECK(d >= 1 && d < L, RR_EINVAL, "order d=%llu must satisfy 1 <= d < L=%llu", (unsigned long long)d,
             (unsigned long long)L);
    RR_CHECK(p != NTT_PRIME || d + 1 <= (1ull << NTT_MAX_LOGN), RR_EINVAL,
             "order d=%llu above the 2^%u-point transform limit", (unsigned long long)d, NTT_MAX_LOGN);
    const uint32_t P = b * (b + 1) / 2, m = 2 * b;
    const uint64_t len = d + 1;
    hipStream_t st;
    RR_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct S {
        hipStream_t s;
        ~S() { (void)hipStreamDestroy(s); }
    } guard{st};
    DevBuf dseq, G, M;
    rr_status s;
    if ((s = dseq.alloc(len * P * 4)) != RR_OK) return s;
    if ((s = G.alloc((size_t)m * b * len * 4)) != RR_OK) return s;
    if ((s = M.alloc((size_t)m * m * len * 4)) != RR_OK) return s;
    for (uint64_t i = 0; i < len * P; i++)
        RR_CHECK(seq[i] < p, RR_EINVAL, "seq[%llu] >= p", (unsigned long long)i);
    RR_HIP(hipMemcpyAsync(dseq.ptr, seq, len * P * 4, hipMemcpyHostToDevice, st));
    if ((s = build_G_device(dseq.as<uint32_t>(), b, len, G.as<uint32_t>(), st)) != RR_OK) return s;
    if ((s = pm_basis_device(p, b, d, G.as<uint32_t>(), M.as<uint32_t>(), delta_out, st)) != RR_OK) return s;
    RR_HIP(hipMemcpyAsync(basis_out, M.ptr, (size_t)m * m * len * 4, hipMemcpyDeviceToHost, st));
    RR_HIP(hipStreamSynchronize(st));
    return RR_OK;
}

rr_status rr_rank_from_basis(const uint32_t *basis, const int64_t *delta, uint32_t b, uint64_t len, uint32_t p,
                             int64_t *rank_out) {
    clear_error();
    RR_CHECK(basis && delta && rank_out, RR_EINVAL, "null argument");
    RR_CHECK(b >= 1 && b <= 64 && len >= 1, RR_EINVAL, "bad sizes");
    rr_status s;
    *rank_out = rank_from_basis_host(basis, delta, b, len, p, &s);
    return s;
}

}  // extern "C"

extern "C" {

// Natural-order transforms in the reference convention (src/ntt.rs:164-211):
// data [count][2^logn], forward with root 31^((p-1)/n), inverse scaled by 1/n.
// The device transform is DIF (bit-reversed spectrum); the permutation to
// natural order is done on the host here, products never need it.
rr_status rr_ntt(uint32_t *data, uint32_t logn, uint32_t count, int inverse) {
    clear_error();
    RR_CHECK(data && logn >= 1 && logn <= NTT_MAX_LOGN && count >= 1, RR_EINVAL, "bad NTT arguments");
    const uint32_t N = 1u << logn, p = NTT_PRIME;
    for (uint64_t i = 0; i < (uint64_t)N * count; i++) RR_CHECK(data[i] < p, RR_EINVAL, "input >= p");
    std::vector<uint32_t> rev(N);
    for (uint32_t i = 0; i < N; i++) {
        uint32_t r = 0, x = i;
        for (uint32_t k = 0; k < logn; k++) { r = (r << 1) | (x & 1); x >>= 1; }
        rev[i] = r;
    }
    hipStream_t st;
    RR_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct S {
        hipStream_t s;
        ~S() { (void)hipStreamDestroy(s); }
    } guard{st};
    NttTables T;
    rr_status s;
    if ((s = T.init(logn, st)) != RR_OK) return s;
    DevBuf a, o;
    size_t bytes = (size_t)N * count * 4;
    if ((s = a.alloc(bytes)) != RR_OK) return s;
    if ((s = o.alloc(bytes)) != RR_OK) return s;
    std::vector<uint32_t> h((size_t)N * count);
    if (!inverse) {
        RR_HIP(hipMemcpyAsync(a.ptr, data, bytes, hipMemcpyHostToDevice, st));
        if ((s = ntt_forward(T, a.as<uint32_t>(), N, 0, N, o.as<uint32_t>(), count, logn, st)) != RR_OK) return s;
        RR_HIP(hipMemcpyAsync(h.data(), o.ptr, bytes, hipMemcpyDeviceToHost, st));
        RR_HIP(hipStreamSynchronize(st));
        for (uint32_t q = 0; q < count; q++)
            for (uint32_t i = 0; i < N; i++) data[(size_t)q * N + rev[i]] = h[(size_t)q * N + i];
    } else {
        for (uint32_t q = 0; q < count; q++)
            for (uint32_t i = 0; i < N; i++) h[(size_t)q * N + i] = data[(size_t)q * N + rev[i]];
        RR_HIP(hipMemcpyAsync(a.ptr, h.data(), bytes, hipMemcpyHostToDevice, st));
        if ((s = ntt_inverse(T, a.as<uint32_t>(), count, logn, 0, N, o.as<uint32_t>(), N, st)) != RR_OK) return s;
        RR_HIP(hipMemcpyAsync(h.data(), o.ptr, bytes, hipMemcpyDeviceToHost, st));
        RR_HIP(hipStreamSynchronize(st));
        uint32_t ninv = h_invmod(N, p);
        for (size_t i = 0; i < (size_t)N * count; i++) data[i] = h_mulmod(h[i], ninv, p);
    }
    return RR_OK;
}

// Truncated polynomial-matrix product (src/poly_mat_mul.rs:89-256):
// C [rows][cols][hi-lo] = coefficients [lo, hi) of A [rows][inner][la] * B [inner][cols][lb], exact mod p.
rr_status rr_poly_mat_mul(const uint32_t *A, uint64_t la, const uint32_t *B, uint64_t lb, uint32_t rows,
                          uint32_t inner, uint32_t cols, uint64_t lo, uint64_t hi, uint32_t p, uint32_t *C) {
    clear_error();
    RR_CHECK(A && B && C && la >= 1 && lb >= 1 && hi > lo && rows && inner && cols, RR_EINVAL,
             "bad product arguments");
    RR_CHECK(inner <= 128 && hi <= la + lb - 1 + 0ull, RR_EINVAL, "inner <= 128 and hi <= la + lb - 1");
    hipStream_t st;
    RR_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct S {
        hipStream_t s;
        ~S() { (void)hipStreamDestroy(s); }
    } guard{st};
    PmCtx c;
    c.p = p;
    c.b = 0;
    c.m = 0;
    c.st = st;
    c.use_ntt = (p == NTT_PRIME);
    DevBuf dA, dB, dC;
    rr_status s;
    size_t sa = (size_t)rows * inner * la, sbn = (size_t)inner * cols * lb, sc = (size_t)rows * cols * (hi - lo);
    if ((s = dA.alloc(sa * 4)) != RR_OK) return s;
    if ((s = dB.alloc(sbn * 4)) != RR_OK) return s;
    if ((s = dC.alloc(sc * 4)) != RR_OK) return s;
    RR_HIP(hipMemcpyAsync(dA.ptr, A, sa * 4, hipMemcpyHostToDevice, st));
    RR_HIP(hipMemcpyAsync(dB.ptr, B, sbn * 4, hipMemcpyHostToDevice, st));
    if (c.use_ntt) {
        uint32_t need = (uint32_t)std::max<uint64_t>(hi, la + lb - 1 - lo), logn = 1;
        while ((1u << logn) < need) logn++;
        RR_CHECK(logn <= NTT_MAX_LOGN, RR_EINVAL, "product too long");
        if ((s = c.tables.init(logn, st)) != RR_OK) return s;
        c.fcap = (size_t)std::max({rows * inner, inner * cols, rows * cols}) << logn;
        if ((s = c.fa.alloc(c.fcap * 4)) != RR_OK) return s;
        if ((s = c.fb.alloc(c.fcap * 4)) != RR_OK) return s;
        if ((s = c.fc.alloc(c.fcap * 4)) != RR_OK) return s;
    }
    if ((s = pmat_mul(c, dA.as<uint32_t>(), (uint32_t)la, dB.as<uint32_t>(), lb, 0, (uint32_t)lb, dC.as<uint32_t>(),
                      rows, inner, cols, (uint32_t)lo, (uint32_t)hi)) != RR_OK)
        return s;
    RR_HIP(hipMemcpyAsync(C, dC.ptr, sc * 4, hipMemcpyDeviceToHost, st));
    RR_HIP(hipStreamSynchronize(st));
    return RR_OK;
}

}  // extern "C"
