/*
 * krylov_oracle.c -- CPU restatement of the reference's Krylov-sequence path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP engine in
 * rust_rank_amd/csrc.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it (oracle/liboracle.so); the product path never
 * links or calls it.
 *
 * What it restates (reference file:line):
 *   orc_transpose      src/matrices.rs:734-778   counting-sort transpose, row order kept
 *   orc_scale_rows     src/matrices.rs:824-840   values[i] = values[i]*d[row] % p
 *   orc_scale_cols     src/matrices.rs:813-821   values[i] = values[i]*d[col] % p
 *   orc_spmm           src/matrices.rs:535-560   y[r] = (sum_e val_e * x[col_e]) % p, b lanes
 *   orc_dot            src/matrices.rs:171-187   chunked dot product mod p
 *   orc_krylov         src/main.rs:88-216 + src/vectorstream.rs:46-90 + src/main.rs:520-526
 *                      a = D_r A D_c, at = D_c A^T, x_{k+1} = at (a x_k),
 *                      S_0 = x_0^T x_0, S_{2k+1} = x_k^T x_{k+1}, S_{2k+2} = x_{k+1}^T x_{k+1},
 *                      upper triangle i<=j in row-major pair order.
 *   orc_dense_rank     src/modular_linalg.rs:132-183  Gaussian elimination rank mod p
 *
 * Arithmetic is widened (unsigned __int128 accumulators) so it stays exact for
 * any p < 2^31; the reference's own u32 lazy reduction is only exact for tiny p
 * (src/matrices.rs:861-870).  Every result is the canonical residue in [0,p),
 * so any exact evaluation order gives identical bits.
 *
 * Parity pins (tests/test_oracle_*.py): dense brute force V^T B^t V at n<=300,
 * the Hankel identity x_a.x_b depends only on a+b, transpose/scale against the
 * reference's compiled C++ (oracle/_ref) on cuda_test/matrix.txt, and the rank
 * against planted deficits.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;

static inline uint32_t mulmod32(uint32_t a, uint32_t b, uint32_t p) {
    return (uint32_t)(((uint64_t)a * b) % p);
}

/* src/matrices.rs:734-778 */
int orc_transpose(uint64_t m, uint64_t n, const uint64_t *rp, const uint32_t *col,
                  const uint32_t *val, uint64_t *rpt, uint32_t *colt, uint32_t *valt) {
    uint64_t nnz = rp[m];
    memset(rpt, 0, (n + 1) * sizeof(uint64_t));
    for (uint64_t e = 0; e < nnz; e++) {
        if (col[e] >= n) return -1;
        rpt[col[e] + 1]++;
    }
    for (uint64_t i = 0; i < n; i++) rpt[i + 1] += rpt[i];
    uint64_t *next = (uint64_t *)malloc(n * sizeof(uint64_t) + 1);
    if (!next) return -2;
    memcpy(next, rpt, n * sizeof(uint64_t));
    for (uint64_t r = 0; r < m; r++) {
        for (uint64_t e = rp[r]; e < rp[r + 1]; e++) {
            uint64_t pos = next[col[e]]++;
            valt[pos] = val[e];
            colt[pos] = (uint32_t)r;
        }
    }
    free(next);
    return 0;
}

/* src/matrices.rs:824-840 */
void orc_scale_rows(uint64_t m, const uint64_t *rp, uint32_t *val, const uint32_t *d, uint32_t p) {
    for (uint64_t r = 0; r < m; r++)
        for (uint64_t e = rp[r]; e < rp[r + 1]; e++) val[e] = mulmod32(val[e], d[r], p);
}

/* src/matrices.rs:813-821 */
void orc_scale_cols(uint64_t nnz, const uint32_t *col, uint32_t *val, const uint32_t *d, uint32_t p) {
    for (uint64_t e = 0; e < nnz; e++) val[e] = mulmod32(val[e], d[col[e]], p);
}

/* src/matrices.rs:535-560: y (m x b, row-interleaved) = A x mod p, x (n x b row-interleaved) */
void orc_spmm(uint64_t m, const uint64_t *rp, const uint32_t *col, const uint32_t *val,
              uint32_t b, const uint32_t *x, uint32_t *y, uint32_t p, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256)
#endif
    for (int64_t r = 0; r < (int64_t)m; r++) {
        u128 acc[64];
        uint32_t bb = b;
        for (uint32_t j = 0; j < bb; j++) acc[j] = 0;
        for (uint64_t e = rp[r]; e < rp[r + 1]; e++) {
            const uint32_t *xr = x + (uint64_t)col[e] * bb;
            uint64_t v = val[e];
            for (uint32_t j = 0; j < bb; j++) acc[j] += (uint64_t)(v * xr[j]);
        }
        for (uint32_t j = 0; j < bb; j++) y[(uint64_t)r * bb + j] = (uint32_t)(acc[j] % p);
    }
}

/* src/matrices.rs:171-187 (any exact order gives the same residue) */
uint32_t orc_dot(uint64_t len, uint64_t sx, const uint32_t *x, uint64_t sy, const uint32_t *y, uint32_t p) {
    u128 acc = 0;
    for (uint64_t i = 0; i < len; i++) acc += (uint64_t)x[i * sx] * y[i * sy];
    return (uint32_t)(acc % p);
}

/* Upper-triangle Gram: out[ii] = sum_r X[r,i] Y[r,j] mod p for i<=j, ii row-major over i<=j
 * (src/main.rs:113-126, 138-156). */
static void gram_upper(uint64_t rows, uint32_t b, const uint32_t *X, const uint32_t *Y,
                       uint32_t p, uint32_t *out, int nthreads) {
    uint32_t P = b * (b + 1) / 2;
    u128 *tot = (u128 *)calloc(P, sizeof(u128));
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
    {
        u128 *loc = (u128 *)calloc(P, sizeof(u128));
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int64_t r = 0; r < (int64_t)rows; r++) {
            const uint32_t *xr = X + (uint64_t)r * b, *yr = Y + (uint64_t)r * b;
            uint32_t ii = 0;
            for (uint32_t i = 0; i < b; i++) {
                uint64_t xi = xr[i];
                for (uint32_t j = i; j < b; j++) loc[ii++] += xi * yr[j];
            }
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        for (uint32_t k = 0; k < P; k++) tot[k] += loc[k] % p;
        free(loc);
    }
    for (uint32_t k = 0; k < P; k++) out[k] = (uint32_t)(tot[k] % p);
    free(tot);
}

/*
 * Krylov sequence following src/main.rs:88-216 with the streams of
 * src/vectorstream.rs:46-90 and preconditioning of src/main.rs:520-526.
 * Input: raw A (CSR, values in [0,p)), D_r (m), D_c (n), x_in = x_k (n x b interleaved).
 * Produces S_t for t in [t_begin, t_end) into seq_out[(t - t_begin) * P + ii],
 * where t_begin is 2k (then S_{2k} = x_k^T x_k is emitted first) or 2k+1.
 * x_out (may be NULL) receives x_K with K = floor((t_end-1)/2).
 */
int orc_krylov(uint64_t m, uint64_t n, const uint64_t *rp, const uint32_t *col, const uint32_t *val,
               uint32_t p, uint32_t b, const uint32_t *dr, const uint32_t *dc,
               const uint32_t *x_in, uint64_t t_begin, uint64_t t_end,
               uint32_t *seq_out, uint32_t *x_out, int nthreads) {
    if (b == 0 || b > 64 || t_end < t_begin) return -1;
    uint64_t nnz = rp[m];
    uint32_t P = b * (b + 1) / 2;
    /* a = D_r A D_c ; at = D_c A^T */
    uint32_t *va = (uint32_t *)malloc(nnz * 4 + 4);
    uint64_t *rpt = (uint64_t *)malloc((n + 1) * 8);
    uint32_t *colt = (uint32_t *)malloc(nnz * 4 + 4);
    uint32_t *vat = (uint32_t *)malloc(nnz * 4 + 4);
    uint32_t *x = (uint32_t *)malloc(n * b * 4 + 4);
    uint32_t *xn = (uint32_t *)malloc(n * b * 4 + 4);
    uint32_t *y = (uint32_t *)malloc(m * b * 4 + 4);
    if (!va || !rpt || !colt || !vat || !x || !xn || !y) return -2;
    if (orc_transpose(m, n, rp, col, val, rpt, colt, vat) != 0) return -3;
    memcpy(va, val, nnz * 4);
    orc_scale_rows(m, rp, va, dr, p);
    orc_scale_cols(nnz, col, va, dc, p);
    orc_scale_rows(n, rpt, vat, dc, p);
    memcpy(x, x_in, n * b * 4);
    uint64_t t = t_begin;
    uint32_t *out = seq_out;
    if (t < t_end && (t % 2) == 0) {
        gram_upper(n, b, x, x, p, out, nthreads);
        out += P;
        t++;
    }
    while (t < t_end) {
        /* t is odd: t = 2k+1 */
        orc_spmm(m, rp, col, va, b, x, y, p, nthreads);
        orc_spmm(n, rpt, colt, vat, b, y, xn, p, nthreads);
        gram_upper(n, b, x, xn, p, out, nthreads);
        out += P;
        t++;
        if (t < t_end) {
            gram_upper(n, b, xn, xn, p, out, nthreads);
            out += P;
            t++;
            uint32_t *tmp = x; x = xn; xn = tmp;
        }
    }
    /* x now holds x_K for the last even index reached */
    if (x_out) memcpy(x_out, x, n * b * 4);
    free(va); free(rpt); free(colt); free(vat); free(x); free(xn); free(y);
    return 0;
}

static uint32_t powmod32(uint32_t a, uint64_t e, uint32_t p) {
    uint64_t r = 1, x = a % p;
    while (e) {
        if (e & 1) r = r * x % p;
        x = x * x % p;
        e >>= 1;
    }
    return (uint32_t)r;
}

/* src/modular_linalg.rs:132-183, dense rows x cols, row-major, destroyed in place. */
uint64_t orc_dense_rank(uint64_t rows, uint64_t cols, uint32_t *mat, uint32_t p, int nthreads) {
    uint64_t rank = 0;
    for (uint64_t i = 0; i < rows * cols; i++) mat[i] %= p;
    for (uint64_t c = 0; c < cols && rank < rows; c++) {
        uint64_t piv = rank;
        while (piv < rows && mat[piv * cols + c] == 0) piv++;
        if (piv == rows) continue;
        if (piv != rank) {
            for (uint64_t j = 0; j < cols; j++) {
                uint32_t t = mat[rank * cols + j];
                mat[rank * cols + j] = mat[piv * cols + j];
                mat[piv * cols + j] = t;
            }
        }
        uint32_t *pr = mat + rank * cols;
        uint32_t inv = powmod32(pr[c], p - 2, p);
        for (uint64_t j = c; j < cols; j++) pr[j] = mulmod32(pr[j], inv, p);
#ifdef _OPENMP
        if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
        for (int64_t r = (int64_t)rank + 1; r < (int64_t)rows; r++) {
            uint32_t *rr = mat + (uint64_t)r * cols;
            uint32_t f = rr[c];
            if (f == 0) continue;
            uint64_t nf = p - f;
            for (uint64_t j = c; j < cols; j++)
                rr[j] = (uint32_t)((rr[j] + nf * pr[j]) % p);
        }
        rank++;
    }
    return rank;
}
