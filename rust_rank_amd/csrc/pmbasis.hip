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
// src/sigma_basis.rs:22-29): a node of order d > K (the job's leaf order) gives
// its left part about half of its ceil(d / K) leaves, all of full order K, so
// every leaf but the last has order K; for powers of two this is the
// reference's split d1 = d2 = d/2.  The basis is the product of the order-1
// steps M_{d-1} ... M_0, each computed from the constant term of the exact
// residual, so any split gives the same bits.
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
#include "pmbasis.hpp"
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include <algorithm>
#include <cstring>

namespace rr {

// ---------------------------------------------------------------------------
// Leaf: up to the job's leaf order of M-Basis steps in one workgroup.  Step s takes the
// constant term of the current residual P = M_s G (coefficient s), sorts its
// rows by delta (stable, descending, as Rust's sort_by), eliminates LSP-style
// (pivot = first nonzero position >= rank in the swapped column order,
// src/modular_linalg.rs:188-239) fraction-free, so no pivot needs an inverse
// inside the loop: every row carries its multiplier sigma_i and its
// coefficients on earlier pivot rows; one parallel inversion gives Linv.
// The step basis is T = D1 Linv pi + x Dx Linv pi (src/sigma_basis.rs:384-396),
// and both P and M advance by it: row i of T X is row perm(i) of X plus the
// Linv coefficients times the pivot rows, shifted by x when row i was a
// pivot row.  This is the reference's chain of order-1 leaves, kept in one
// launch.
// Leaves per launch.  The pack (12 KB) travels in the kernel arguments;
// fewer, fuller launches per round took a 512-matrix batch from 3.23 to
// 3.12 s (with NTT_PACK 96 instead of 32).
constexpr int LEAF_PACK = 256;
struct LeafPack {
    LeafDesc d[LEAF_PACK];
};

// LDS of a leaf: the residual P [m][n][K] and the basis M [m][m][K+1]
// stay resident (K > 1), beside the elimination arrays R, U, the row
// multipliers and permutations and the row degrees (delta).
struct LeafLds {
    uint32_t *R, *U, *sig, *fcol, *perm, *pivrow, *isx;
    uint32_t *prow;  // per pivot q: the row of X it reads, perm(pivrow q)
    uint32_t *pbuf;  // two slots of LEAF_PBUF words: a pivot row and its pivot fields
    uint32_t *cs;    // per row of P and M (K > 1): the factor c it is kept scaled by
    uint32_t *P, *M;
    int64_t *dl, *dln;
};

constexpr uint32_t LEAF_PBUF = 64 * 2 + 8;  // a row of up to 128 columns, then {position, column, entry, sigma}
__host__ __device__ constexpr uint32_t leaf_dl_off(uint32_t m, uint32_t n) {
    return (2 * m * n + 5 * m + 2 * n + 2 * LEAF_PBUF + 1) & ~1u;
}

__device__ __forceinline__ LeafLds leaf_lds(uint32_t m, uint32_t n, uint32_t K) {
    extern __shared__ __align__(16) uint32_t sm[];
    LeafLds L;
    L.R = sm;
    L.U = L.R + m * n;
    L.sig = L.U + m * n;
    L.fcol = L.sig + m;
    L.perm = L.fcol + m;
    L.isx = L.perm + m;
    L.pivrow = L.isx + m;
    L.prow = L.pivrow + n;
    L.pbuf = L.prow + n;
    L.cs = L.pbuf + 2 * LEAF_PBUF;
    L.dl = (int64_t *)(sm + leaf_dl_off(m, n));  // 8-byte aligned
    L.dln = L.dl + m;
    L.P = (uint32_t *)(L.dln + m);
    L.M = L.P + (K > 1 ? m * n * K : 0);
    return L;
}

static size_t leaf_lds_words(uint32_t m, uint32_t n, uint32_t K) {
    size_t w = (size_t)leaf_dl_off(m, n) + 4 * m;
    if (K > 1) w += (size_t)m * n * K + (size_t)m * m * (K + 1);
    return w;
}

constexpr uint32_t LEAF_THREADS = 1024;
// Advance tiles: a wavefront takes ADV_ROWS rows x 64 consecutive coefficient
// slots of X, so each pivot row's word is loaded once for ADV_ROWS rows and
// each U coefficient is one broadcast read; a wavefront holds at most
// ADV_ITEMS such tiles.
constexpr uint32_t ADV_ROWS = 4, ADV_ITEMS = 5;
__host__ __device__ constexpr uint32_t adv_items(uint32_t m, uint32_t cols, uint32_t nsl) {
    return ((cols * nsl + 63) / 64) * ((m + ADV_ROWS - 1) / ADV_ROWS);
}

// In-place X <- T X on coefficients: X [m][cols][stride] valid on [klo, khi],
// rewritten on [olo, ohi].  Row i of T X is row perm(i) of X plus the Linv
// coefficients U'[i][q] times the pivot rows L.prow[q], shifted up one
// coefficient when row i was a pivot row.  Every output is computed into
// registers before any is stored (inputs and outputs share X); outputs no
// input reaches are zeroed.  Tile (row group, slot block) of ADV_ROWS rows and
// 64 slots (slot = col * nsl + coefficient - klo, lanes consecutive, so a
// wavefront reads runs of consecutive words of X): tile t belongs to
// wavefront t mod W.  A coefficient u enters as u_hi 2^16 + u_lo with 64-bit
// integer multiply-adds; with at most 64 pivots both sums stay below 2^53,
// so they convert to fp64 exactly.
__device__ __forceinline__ void leaf_advance(const LeafLds &L, uint32_t *X, uint32_t cols, uint32_t stride, uint32_t klo,
                                             uint32_t khi, uint32_t olo, uint32_t ohi, uint32_t rank, uint32_t m,
                                             uint32_t n, uint32_t p, double pd, double inv) {
    const uint32_t tid = threadIdx.x, lane = tid & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6), nw = LEAF_THREADS / 64;
    const uint32_t nsl = khi - klo + 1, ncc = cols * nsl, ccb = (ncc + 63) / 64;
    const uint32_t items = ccb * ((m + ADV_ROWS - 1) / ADV_ROWS), rstride = cols * stride;
    // lane q holds pivot row q's offset in X (rank <= n <= 64)
    const uint32_t prow_l = lane < rank ? L.prow[lane] * rstride : 0u;
    uint32_t res[ADV_ITEMS][ADV_ROWS];
#pragma unroll
    for (uint32_t it = 0; it < ADV_ITEMS; it++) {
        const uint32_t item = wave + it * nw;
        if (item >= items) break;
        const uint32_t i0 = (item / ccb) * ADV_ROWS, cc = (item % ccb) * 64 + lane;
        const uint32_t ccc = min(cc, ncc - 1), col = ccc / nsl, xo = col * stride + klo + (ccc - col * nsl);
        uint32_t uo[ADV_ROWS];
        uint64_t hacc[ADV_ROWS], lacc[ADV_ROWS];
#pragma unroll
        for (uint32_t r = 0; r < ADV_ROWS; r++) {
            const uint32_t i = min(i0 + r, m - 1);
            uo[r] = i * n;
            hacc[r] = 0;
            lacc[r] = mulmod(X[L.perm[i] * rstride + xo], L.sig[i], p, inv);
        }
#pragma unroll 2
        for (uint32_t q = 0; q < rank; q++) {
            const uint32_t x = X[__builtin_amdgcn_readlane(prow_l, q) + xo];
#pragma unroll
            for (uint32_t r = 0; r < ADV_ROWS; r++) {
                const uint32_t u = L.U[uo[r] + q];
                hacc[r] += (uint64_t)(u >> 16) * x;
                lacc[r] += (uint64_t)(u & 0xFFFFu) * x;
            }
        }
#pragma unroll
        for (uint32_t r = 0; r < ADV_ROWS; r++)
            res[it][r] =
                (uint32_t)fold(fma(fold((double)hacc[r], pd, inv), 65536.0, fold((double)lacc[r], pd, inv)), pd, inv);
    }
    __syncthreads();
#pragma unroll
    for (uint32_t it = 0; it < ADV_ITEMS; it++) {
        const uint32_t item = wave + it * nw;
        if (item >= items) break;
        const uint32_t i0 = (item / ccb) * ADV_ROWS, cc = (item % ccb) * 64 + lane;
        if (cc >= ncc) continue;
        const uint32_t col = cc / nsl, kk = klo + (cc - col * nsl);
#pragma unroll
        for (uint32_t r = 0; r < ADV_ROWS; r++) {
            const uint32_t i = i0 + r;
            if (i >= m) break;
            const uint32_t ko = kk + L.isx[i];
            if (ko >= olo && ko <= ohi) X[i * rstride + col * stride + ko] = res[it][r];
        }
    }
    for (uint32_t idx = tid; idx < m * cols; idx += blockDim.x) {
        const uint32_t sh = L.isx[idx / cols];
        for (uint32_t ko = olo; ko <= ohi; ko++)
            if (ko < sh || ko - sh < klo || ko - sh > khi) X[idx * stride + ko] = 0u;
    }
    __syncthreads();
}

// One M-Basis step on the constant term already in L.R (rows in the delta
// order L.perm): fraction-free LSP, then U' = Linv coefficients, isx, the
// new delta in L.dln.  Returns the rank.
//
// A row is [R | U] (n + n columns) and its multiplier sigma: U holds its
// coefficients on the pivot rows.  One update of row i by the pivot of row k
// (column c, pivot index q):
//   x <- (pkv x + (p - f) y) 2^-32,  pkv = R[k][c], f = R[i][c],
// y = row k's column, except U column q, where row i holds 0 and y is
// sigma_k; sigma_i <- pkv sigma_i 2^-32.  Scaling a whole row with its
// multiplier (the Montgomery factor, and pkv when f = 0) leaves its Linv row
// U / sigma and its zero pattern as they are; pkv x + (p - f) y < 2 p^2.
//
// The rows stay in registers for the whole elimination: row i belongs to
// wavefront i mod W (W wavefronts), slot i / W; lanes run over the columns
// (column j = 64 g + lane in register g, NREG = 1 for n <= 32, 2 for
// n <= 64), sigma is a uniform register.  For each row k in order its owner
// searches it (the reference's rule: the first position >= rank, in a column
// order every wavefront keeps in one register, whose column is nonzero),
// then writes the pivot row and its pivot to one of two LDS slots; after one
// barrier every wavefront applies the pivot to its rows below k.  Only the
// pivot row crosses LDS, once; one barrier a row.
// Montgomery reduction with R = 2^32: x < p 2^32 -> x R^-1 mod p; pneg = -p^-1 mod 2^32.
__device__ __forceinline__ uint32_t redc(uint64_t x, uint32_t p, uint32_t pneg) {
    const uint32_t mq = (uint32_t)x * pneg;
    const uint32_t r = (uint32_t)((x + (uint64_t)mq * p) >> 32);  // < 2p
    return r >= p ? r - p : r;
}

__device__ __forceinline__ uint32_t neg_inv_2_32(uint32_t p) {  // p odd
    uint32_t x = p;  // correct to 3 bits; each Newton step doubles them
#pragma unroll
    for (int i = 0; i < 4; i++) x *= 2u - p * x;
    return 0u - x;
}

template <int NREG, int S>
__device__ __forceinline__ uint32_t leaf_lsp(const LeafLds &L, uint32_t m, uint32_t n, uint32_t p) {
    const uint32_t tid = threadIdx.x, lane = tid & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6), nw = LEAF_THREADS / 64;
    const uint32_t pneg = neg_inv_2_32(p);
    uint32_t x[S][NREG], sg[S];
#pragma unroll
    for (int s = 0; s < S; s++) {
        const uint32_t i = s * nw + wave, ic = i < m ? i : m - 1;
#pragma unroll
        for (int g = 0; g < NREG; g++) {
            const uint32_t j = 64 * g + lane;
            const uint32_t vr = L.R[ic * n + min(j, n - 1)], vu = L.U[ic * n + min(j - n, n - 1)];
            x[s][g] = j < n ? vr : j < 2 * n ? vu : 0u;
        }
        sg[s] = L.sig[ic];
    }
    uint32_t cord = lane;  // lane q: the column at position q of the column order
    uint32_t rank = 0;
    for (uint32_t k = 0; k < m && rank < n; k++) {  // rank n: every later row is zero
        uint32_t *B = L.pbuf + (k & 1) * LEAF_PBUF;
        if (wave == k % nw) {
            const uint32_t ks = k / nw;
            uint32_t r0 = x[0][0], rs = sg[0];
#pragma unroll
            for (int s = 1; s < S; s++) {
                r0 = (uint32_t)s == ks ? x[s][0] : r0;
                rs = (uint32_t)s == ks ? sg[s] : rs;
            }
            // lane q: row k's entry in the column at position q
            const uint32_t v = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(min(cord, 63u) * 4), (int)r0);
            const uint64_t bal = __ballot(lane >= rank && lane < n && v != 0);
            const uint32_t pos = bal ? (uint32_t)__ffsll((unsigned long long)bal) - 1 : n;
            if (pos < n) {
#pragma unroll
                for (int g = 0; g < NREG; g++) {
                    uint32_t rg = x[0][g];
#pragma unroll
                    for (int s = 1; s < S; s++) rg = (uint32_t)s == ks ? x[s][g] : rg;
                    B[64 * g + lane] = rg;
                }
            }
            // cross-lane reads stay outside the lane-0 branch: inside it the
            // compiler may replace their operand by its value in lane 0
            const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)cord, (int)(pos < n ? pos : 0));
            const uint32_t pkv = (uint32_t)__builtin_amdgcn_readlane((int)r0, (int)c);
            if (lane == 0) {
                B[128] = pos;
                if (pos < n) {
                    B[129] = c;
                    B[130] = pkv;
                    B[131] = rs;
                    L.pivrow[rank] = k;
                }
            }
        }
        __syncthreads();
        // all reads of the slot issued together (the row is stale when row k has no pivot; unused then)
        const uint32_t pos = __builtin_amdgcn_readfirstlane(B[128]);
        const uint32_t c = __builtin_amdgcn_readfirstlane(B[129]);
        const uint32_t pkv = __builtin_amdgcn_readfirstlane(B[130]);
        const uint32_t sk = __builtin_amdgcn_readfirstlane(B[131]);
        uint32_t y[NREG];
#pragma unroll
        for (int g = 0; g < NREG; g++) y[g] = B[64 * g + lane];
        if (pos >= n) continue;  // row k is zero past the pivots: no pivot
        // swap positions pos and rank of the column order
        const uint32_t cr = (uint32_t)__builtin_amdgcn_readlane((int)cord, (int)rank);
        if (lane == pos) cord = cr;
        if (lane == rank) cord = c;
        const uint32_t q = rank++;
#pragma unroll
        for (int g = 0; g < NREG; g++) y[g] = 64 * g + lane == n + q ? sk : y[g];
#pragma unroll
        for (int s = 0; s < S; s++) {
            const uint32_t i = s * nw + wave;
            if (i > k && i < m) {
                const uint32_t nf = p - (uint32_t)__builtin_amdgcn_readlane((int)x[s][0], (int)c);
#pragma unroll
                for (int g = 0; g < NREG; g++) x[s][g] = redc((uint64_t)pkv * x[s][g] + (uint64_t)nf * y[g], p, pneg);
                sg[s] = redc((uint64_t)pkv * sg[s], p, pneg);
            }
        }
    }
#pragma unroll
    for (int s = 0; s < S; s++) {
        const uint32_t i = s * nw + wave;
        if (i < m) {
#pragma unroll
            for (int g = 0; g < NREG; g++) {
                const uint32_t j = 64 * g + lane;
                if (j < n) L.R[i * n + j] = x[s][g];
                else if (j < 2 * n) L.U[i * n + j - n] = x[s][g];
            }
            if (lane == 0) L.sig[i] = sg[s];
        }
    }
    __syncthreads();
    return rank;
}

// Rows per wavefront and registers per row for m rows of n columns (m <= 128, n <= 64).
__device__ __forceinline__ uint32_t leaf_lsp_any(const LeafLds &L, uint32_t m, uint32_t n, uint32_t p) {
    const uint32_t nw = LEAF_THREADS / 64;
    if (n > 32) return leaf_lsp<2, 128 / 16>(L, m, n, p);
    if (m <= nw) return leaf_lsp<1, 1>(L, m, n, p);
    if (m <= 2 * nw) return leaf_lsp<1, 2>(L, m, n, p);
    return leaf_lsp<1, 4>(L, m, n, p);
}

// M <- T M for M = I (a multi-step leaf's first step): row i of T is
// sigma_i at column perm(i) plus U_iq at column prow(q), and no two of these
// meet (prow(q) = perm(i) only for q = row i's own pivot, where U_iq = 0), so
// the product is T itself, each row at coefficient isx_i, with no arithmetic.
// L.R holds the elimination's rows, free by now: it maps a row of M to its
// pivot index + 1 (0: none).
__device__ __forceinline__ void leaf_advance_identity(const LeafLds &L, uint32_t *M, uint32_t m, uint32_t n,
                                                      uint32_t MS, uint32_t rank) {
    const uint32_t tid = threadIdx.x;
    for (uint32_t j = tid; j < m; j += LEAF_THREADS) L.R[j] = 0;
    __syncthreads();
    for (uint32_t q = tid; q < rank; q += LEAF_THREADS) L.R[L.prow[q]] = q + 1;
    __syncthreads();
    for (uint32_t idx = tid; idx < m * m; idx += LEAF_THREADS) {
        const uint32_t i = idx / m, j = idx - i * m, qq = L.R[j], sh = L.isx[i];
        const uint32_t v = j == L.perm[i] ? L.sig[i] : qq ? L.U[i * n + qq - 1] : 0u;
        M[(size_t)idx * MS + sh] = v;
        M[(size_t)idx * MS + 1 - sh] = 0u;
    }
    __syncthreads();
}

// The elimination and what the advance needs from it.  normalize: Linv =
// U / sigma (one inversion per row; a single-step leaf writes its basis from
// it).  Otherwise rows stay scaled: the step basis of the scaled rows is row
// i = sigma_i e_perm(i) + U_i (the fraction-free result itself), which is
// sigma_i c_perm(i) times the true step applied to the true rows when row r
// is kept as c_r times the true one; the new factors go to fcol.
__device__ __forceinline__ uint32_t leaf_eliminate(const LeafLds &L, uint32_t m, uint32_t n, uint32_t p, double inv,
                                                   bool normalize) {
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    const uint32_t rank = leaf_lsp_any(L, m, n, p);
    for (uint32_t i = tid; i < m; i += nt) {
        if (normalize) {  // Linv row i = e_i + (U row i) / sigma_i over the earlier pivot rows
            uint32_t sv = L.sig[i], r = 1, e = p - 2;
            while (e) {
                if (e & 1) r = mulmod(r, sv, p, inv);
                sv = mulmod(sv, sv, p, inv);
                e >>= 1;
            }
            L.fcol[i] = r;
        } else {
            L.fcol[i] = mulmod(L.sig[i], L.cs[L.perm[i]], p, inv);
        }
    }
    for (uint32_t q = tid; q < rank; q += nt) {
        L.isx[L.pivrow[q]] = 1;
        L.prow[q] = L.perm[L.pivrow[q]];
    }
    __syncthreads();
    // U[i][q] is nonzero only for pivots of earlier rows (pivrow q < i): only they update row i
    if (normalize)
        for (uint32_t idx = tid; idx < m * rank; idx += nt) {
            const uint32_t i = idx / rank, q = idx % rank;
            L.U[i * n + q] = mulmod(L.U[i * n + q], L.fcol[i], p, inv);
        }
    for (uint32_t i = tid; i < m; i += nt) L.dln[i] = L.dl[L.perm[i]] - (int64_t)L.isx[i];
    __syncthreads();
    return rank;
}

// Sorts rows by delta (stable, descending: rank_i = #{j: dl_j > dl_i} +
// #{j < i: dl_j == dl_i}) and resets the elimination state.  A wavefront
// ranks one row at a time: lane j compares row j with it, one ballot counts.
__device__ __forceinline__ void leaf_sort(const LeafLds &L, uint32_t m) {
    const uint32_t tid = threadIdx.x, lane = tid & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6), nw = LEAF_THREADS / 64;
    const int64_t d0 = L.dl[min(lane, m - 1)], d1 = L.dl[min(64 + lane, m - 1)];
    for (uint32_t i = wave; i < m; i += nw) {
        const int64_t di = L.dl[i];
        uint32_t r = (uint32_t)__popcll(__ballot(lane < m && (d0 > di || (d0 == di && lane < i))));
        if (m > 64) r += (uint32_t)__popcll(__ballot(64 + lane < m && (d1 > di || (d1 == di && 64 + lane < i))));
        if (lane == 0) L.perm[r] = i;
    }
    for (uint32_t i = tid; i < m; i += blockDim.x) {
        L.sig[i] = 1;
        L.isx[i] = 0;
    }
    __syncthreads();
}

// D.G: residual window G(i, j, k) = G[(i * n + j) * gs + k], k < d; D.M: [m][m][d + 1].
__global__ __launch_bounds__(LEAF_THREADS) void k_mbasis_iter(LeafPack pk, uint32_t p, double pd, double inv) {
    const LeafDesc D = pk.d[blockIdx.x];
    const uint32_t m = D.m, n = D.n, K = D.d;
    const LeafLds L = leaf_lds(m, n, K);
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    for (uint32_t i = tid; i < m; i += nt) L.dl[i] = D.delta[i];
    __syncthreads();
    if (K == 1) {  // a single step: M = D1 Linv pi + x Dx Linv pi, written directly
        leaf_sort(L, m);
        for (uint32_t idx = tid; idx < m * n; idx += nt) {
            const uint32_t i = idx / n, j = idx % n;
            L.R[idx] = D.G[((size_t)L.perm[i] * n + j) * D.gs] % p;
            L.U[idx] = 0;
        }
        __syncthreads();
        const uint32_t rank = leaf_eliminate(L, m, n, p, inv, true);
        for (uint32_t idx = tid; idx < m * m; idx += nt) {
            D.M[(size_t)idx * 2] = 0;
            D.M[(size_t)idx * 2 + 1] = 0;
        }
        __syncthreads();
        for (uint32_t i = tid; i < m; i += nt) {
            const uint32_t slot = L.isx[i];
            D.M[((size_t)i * m + L.perm[i]) * 2 + slot] = 1;
            for (uint32_t q = 0; q < rank; q++)
                if (L.pivrow[q] < i) D.M[((size_t)i * m + L.perm[L.pivrow[q]]) * 2 + slot] = L.U[i * n + q];
            D.delta[i] = L.dln[i];
        }
        return;
    }
    const uint32_t MS = K + 1;
    for (uint32_t idx = tid; idx < m * n * K; idx += nt) {
        const uint32_t ij = idx / K, k = idx - ij * K;
        L.P[idx] = D.G[(size_t)ij * D.gs + k] % p;
    }
    for (uint32_t idx = tid; idx < m * m * MS; idx += nt) {
        const uint32_t ij = idx / MS, k = idx - ij * MS;
        L.M[idx] = (k == 0 && ij / m == ij % m) ? 1u : 0u;
    }
    for (uint32_t i = tid; i < m; i += nt) L.cs[i] = 1;
    __syncthreads();
    for (uint32_t s = 0; s < K; s++) {
        leaf_sort(L, m);
        for (uint32_t idx = tid; idx < m * n; idx += nt) {
            const uint32_t i = idx / n, j = idx - i * n;
            L.R[idx] = L.P[(L.perm[i] * n + j) * K + s];
            L.U[idx] = 0;
        }
        __syncthreads();
        const uint32_t rank = leaf_eliminate(L, m, n, p, inv, false);
        // P <- T P on the coefficients later steps read (s .. K-1 in, s+1 .. K-1 out)
        if (s + 1 < K) leaf_advance(L, L.P, n, K, s, K - 1, s + 1, K - 1, rank, m, n, p, pd, inv);
        // M <- T M (degree s -> s + 1)
        if (s == 0) leaf_advance_identity(L, L.M, m, n, MS, rank);
        else leaf_advance(L, L.M, m, MS, 0, s, 0, s + 1, rank, m, n, p, pd, inv);
        for (uint32_t i = tid; i < m; i += nt) {
            L.dl[i] = L.dln[i];
            L.cs[i] = L.fcol[i];
        }
        __syncthreads();
    }
    // the true basis: row i of M divided by its factor
    for (uint32_t i = tid; i < m; i += nt) {
        uint32_t sv = L.cs[i], r = 1, e = p - 2;
        while (e) {
            if (e & 1) r = mulmod(r, sv, p, inv);
            sv = mulmod(sv, sv, p, inv);
            e >>= 1;
        }
        L.fcol[i] = r;
    }
    __syncthreads();
    for (uint32_t idx = tid; idx < m * m * MS; idx += nt) D.M[idx] = mulmod(L.M[idx], L.fcol[idx / (m * MS)], p, inv);
    for (uint32_t i = tid; i < m; i += nt) D.delta[i] = L.dl[i];
}

constexpr int MUL_PACK = 48;
struct MulPack {
    MulDesc d[MUL_PACK];
};

// direct products, blockIdx.y = descriptor; thread per output coefficient
__global__ __launch_bounds__(256) void k_pmul_school_batch(MulPack pk, uint32_t p, double pd, double inv) {
    const MulDesc &D = pk.d[blockIdx.y];
    const uint32_t outlen = D.hi - D.lo;
    const uint64_t total = (uint64_t)D.rows * D.cols * outlen;
    for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t k = (uint32_t)(idx % outlen) + D.lo;
        uint64_t ij = idx / outlen;
        uint32_t j = (uint32_t)(ij % D.cols), i = (uint32_t)(ij / D.cols);
        uint32_t s_lo = k >= D.lb ? k - D.lb + 1 : 0, s_hi = min(k + 1, D.la);
        double hacc = 0.0, lacc = 0.0;
        uint32_t terms = 0;
        for (uint32_t r = 0; r < D.inner; r++) {
            const uint32_t *ar = D.A + ((size_t)i * D.inner + r) * D.la;
            const uint32_t *br = D.B + ((size_t)r * D.cols + j) * D.sb + D.ob;
            for (uint32_t s = s_lo; s < s_hi; s++) {
                uint32_t a = ar[s];
                double bd = (double)br[k - s];
                hacc = fma((double)(a >> 16), bd, hacc);
                lacc = fma((double)(a & 0xFFFFu), bd, lacc);
                if (++terms == 64) {
                    hacc = fold(hacc, pd, inv);
                    lacc = fold(lacc, pd, inv);
                    terms = 0;
                }
            }
        }
        double v = fma(fold(hacc, pd, inv), 65536.0, fold(lacc, pd, inv));
        D.C[(size_t)(i * D.cols + j) * outlen + (k - D.lo)] = (uint32_t)fold(v, pd, inv);
    }
}

// Short products with operands staged in LDS: block (x, i, desc) computes row
// i, columns [x jb, x jb + jb) and all output coefficients.  Only the A
// coefficients below hi and the B coefficients that meet [lo, hi) are
// staged.  Each output's sum over the inner index is split over up to
// 256 / outputs threads (residues mod p, added up through LDS), so a block
// keeps all its lanes busy however few outputs it has.
constexpr uint32_t SCH_LDS_WORDS = 12288;
constexpr uint32_t SCH_JB = 4;
struct SchoolTile {
    uint32_t la, t0, lbw, jb;  // staged A length, first staged B coefficient, staged B length, columns per block
};
__host__ __device__ inline SchoolTile school_tile(const MulDesc &D) {
    SchoolTile T;
    T.la = D.la < D.hi ? D.la : D.hi;
    T.t0 = D.lo + 1 > T.la ? D.lo + 1 - T.la : 0;
    const uint32_t t1 = D.hi < D.lb ? D.hi : D.lb;
    T.lbw = t1 > T.t0 ? t1 - T.t0 : 0;
    const uint32_t outlen = D.hi - D.lo;
    const uint32_t fixed = D.inner * T.la + 256 + SCH_JB * outlen;  // A, partial sums
    const uint32_t per_j = D.inner * (T.lbw > 0 ? T.lbw : 1);
    uint32_t jb = fixed < SCH_LDS_WORDS ? (SCH_LDS_WORDS - fixed) / per_j : 0;
    jb = jb < SCH_JB ? jb : SCH_JB;
    T.jb = T.la <= 64 ? (jb < D.cols ? jb : D.cols) : 0;  // 0: does not fit
    return T;
}

__global__ __launch_bounds__(256) void k_pmul_school_lds(MulPack pk, uint32_t p, double pd, double inv) {
    __shared__ uint32_t sm[SCH_LDS_WORDS];
    const MulDesc &D = pk.d[blockIdx.z];
    const uint32_t i = blockIdx.y;
    if (i >= D.rows) return;
    const SchoolTile T = school_tile(D);
    const uint32_t j0 = blockIdx.x * T.jb;
    if (j0 >= D.cols) return;
    const uint32_t jn = min(T.jb, D.cols - j0), outlen = D.hi - D.lo, inner = D.inner, la = T.la, lbw = T.lbw;
    const uint32_t nout = jn * outlen, rs = max(1u, min(inner, 256u / nout));  // inner-index groups
    uint32_t *As = sm, *Bs = sm + inner * la, *part = Bs + inner * jn * lbw;
    for (uint32_t idx = threadIdx.x; idx < inner * la; idx += 256) {
        const uint32_t r = idx / la, sx = idx - r * la;
        As[idx] = D.A[((size_t)i * inner + r) * D.la + sx];
    }
    for (uint32_t idx = threadIdx.x; idx < inner * jn * lbw; idx += 256) {
        const uint32_t rj = idx / lbw, t = idx - rj * lbw, r = rj / jn, jj = rj - r * jn;
        Bs[idx] = D.B[((size_t)r * D.cols + j0 + jj) * D.sb + D.ob + T.t0 + t];
    }
    __syncthreads();
    for (uint32_t w = threadIdx.x; w < nout * rs; w += 256) {
        const uint32_t g = w / nout, o = w - g * nout, jj = o / outlen, k = D.lo + (o - jj * outlen);
        const uint32_t s_lo = k >= D.lb ? k - D.lb + 1 : 0, s_hi = min(k + 1, la);
        double hacc = 0.0, lacc = 0.0;
        uint32_t terms = 0;
        if (s_lo < s_hi) {
            const uint32_t *br = Bs + jj * lbw + (k - T.t0);  // B coefficient k - s of (r, jj) at br[r jn lbw - s]
            for (uint32_t r = g; r < inner; r += rs) {
                if (terms + la > 64) {
                    hacc = fold(hacc, pd, inv);
                    lacc = fold(lacc, pd, inv);
                    terms = 0;
                }
                terms += s_hi - s_lo;
                const uint32_t *ar = As + r * la, *brr = br + r * jn * lbw;
                for (uint32_t sx = s_lo; sx < s_hi; sx++) {
                    const uint32_t a = ar[sx];
                    const double bd = (double)brr[-(int)sx];
                    hacc = fma((double)(a >> 16), bd, hacc);
                    lacc = fma((double)(a & 0xFFFFu), bd, lacc);
                }
            }
        }
        const uint32_t v = (uint32_t)fold(fma(fold(hacc, pd, inv), 65536.0, fold(lacc, pd, inv)), pd, inv);
        if (rs == 1) D.C[(size_t)(i * D.cols + j0 + jj) * outlen + (k - D.lo)] = v;
        else part[w] = v;
    }
    if (rs == 1) return;
    __syncthreads();
    for (uint32_t o = threadIdx.x; o < nout; o += 256) {
        uint32_t v = part[o];
        for (uint32_t g = 1; g < rs; g++) v = addmod(v, part[g * nout + o], p);
        const uint32_t jj = o / outlen, k = o - jj * outlen;
        D.C[(size_t)(i * D.cols + j0 + jj) * outlen + k] = v;
    }
}

// Launches the short products: LDS-tiled where the operands fit, otherwise
// the thread-per-output kernel.
static rr_status launch_school(const std::vector<MulDesc> &ops, uint32_t p, hipStream_t st, uint64_t *launches) {
    const ModP mp = ModP::make(p);
    std::vector<MulDesc> tiled, direct;
    for (const MulDesc &D : ops) (school_tile(D).jb ? tiled : direct).push_back(D);
    for (size_t g = 0; g < tiled.size(); g += MUL_PACK) {
        MulPack pk;
        const uint32_t cnt = (uint32_t)std::min<size_t>(MUL_PACK, tiled.size() - g);
        uint32_t bx = 1, by = 1;
        for (uint32_t k = 0; k < cnt; k++) {
            pk.d[k] = tiled[g + k];
            const SchoolTile T = school_tile(pk.d[k]);
            bx = std::max(bx, (pk.d[k].cols + T.jb - 1) / T.jb);
            by = std::max(by, pk.d[k].rows);
        }
        hipLaunchKernelGGL(k_pmul_school_lds, dim3(bx, by, cnt), dim3(256), 0, st, pk, p, mp.pd, mp.inv);
        RR_HIP(hipGetLastError());
        if (launches) (*launches)++;
    }
    for (size_t g = 0; g < direct.size(); g += MUL_PACK) {
        MulPack pk;
        const uint32_t cnt = (uint32_t)std::min<size_t>(MUL_PACK, direct.size() - g);
        uint64_t maxout = 0;
        for (uint32_t k = 0; k < cnt; k++) {
            pk.d[k] = direct[g + k];
            const MulDesc &D = pk.d[k];
            maxout = std::max<uint64_t>(maxout, (uint64_t)D.rows * D.cols * (D.hi - D.lo));
        }
        const uint32_t bx = (uint32_t)std::min<uint64_t>((maxout + 255) / 256, 4096);
        hipLaunchKernelGGL(k_pmul_school_batch, dim3(std::max(bx, 1u), cnt), dim3(256), 0, st, pk, p, mp.pd, mp.inv);
        RR_HIP(hipGetLastError());
        if (launches) (*launches)++;
    }
    return RR_OK;
}

// ---------------------------------------------------------------------------
// The PM-basis of one matrix as an explicit state machine over the
// recursion of src/sigma_basis.rs:190-219, so that the next device operation
// of many matrices can be launched together.

// Order of the left part of a node of order d > K: K times half (rounded
// up) of its ceil(d / K) leaves; 1 <= d1 < d.
static inline uint64_t split_left(uint64_t d, uint64_t K) {
    const uint64_t leaves = (d + K - 1) / K;
    return K * ((leaves + 1) / 2);
}

// Workspace words along the deepest path of the recursion below an order-d
// node with leaves of order K.
// Transform size of both products of an order-d node (the middle product
// M1 G and M2 M1 both need d + 1 coefficients), and whether the node keeps
// M1's spectrum for the second: only where the first product is an NTT one
// in any round (la lb above the direct-product limits).
constexpr size_t PM_MAX_DEPTH = 64;  // recursion frames per job (orders below 2^20 need about 20)

static uint32_t node_logn(uint64_t d) {
    uint32_t l = 1;
    while ((1ull << l) < d + 1) l++;
    return l;
}
static bool node_keeps(uint64_t d, uint64_t d1) { return (d1 + 1) * (d + 1) > 256 && node_logn(d) <= NTT_MAX_LOGN; }

static uint64_t rec_words(uint64_t d, uint64_t K, uint32_t m, uint32_t b) {
    if (d <= K) return 0;
    const uint64_t d1 = split_left(d, K), d2 = d - d1;
    uint64_t here = ((uint64_t)m * m * (d1 + 1) + 63) + ((uint64_t)m * b * (d2 + 1) + 63) +
                    ((uint64_t)m * m * (d2 + 1) + 63);
    if (node_keeps(d, d1)) here += ((uint64_t)m * m << node_logn(d)) + 63;
    return here + std::max(rec_words(d1, K, m, b), rec_words(d2, K, m, b));
}

struct Op {
    int kind;  // 0 leaf, 1 product
    uint64_t need = 0;  // leading coefficients of the job's G it reads
    Frame *keep = nullptr, *given = nullptr;  // spectrum reuse (products)
    LeafDesc leaf;
    MulDesc mul;
};

// Advances the job to its next device operation; returns false when finished.
static bool next_op(PmJob &J, Op &op, rr_status &err) {
    err = RR_OK;
    while (!J.stack.empty()) {
        size_t fi = J.stack.size() - 1;
        Frame &F = J.stack[fi];
        if (F.d <= J.K) {
            if (F.stage == 0) {
                F.stage = 1;
                op.kind = 0;
                op.leaf = LeafDesc{F.G, F.gs, J.delta, F.M, J.m, J.b, (uint32_t)F.d};
                op.need = F.G == J.G ? F.d + 1 : 0;
                return true;
            }
            J.stack.pop_back();
            continue;
        }
        const uint64_t d1 = split_left(F.d, J.K), d2 = F.d - d1;
        switch (F.stage) {
            case 0: {
                F.mark = J.pool.top;
                F.M1 = J.pool.get((size_t)J.m * J.m * (d1 + 1));
                F.GG = J.pool.get((size_t)J.m * J.b * (d2 + 1));
                F.M2 = J.pool.get((size_t)J.m * J.m * (d2 + 1));
                F.SP = node_keeps(F.d, d1) ? J.pool.get((size_t)J.m * J.m << node_logn(F.d)) : nullptr;
                F.sp_logn = -1;
                if (!F.M1 || !F.GG || !F.M2 || (node_keeps(F.d, d1) && !F.SP)) {
                    set_error("PM-basis workspace exhausted");
                    err = RR_ENOMEM;
                    return false;
                }
                F.stage = 1;
                if (J.stack.size() >= J.stack.capacity()) {
                    set_error("PM-basis recursion deeper than %zu levels", J.stack.capacity());
                    err = RR_EINVAL;
                    return false;
                }
                Frame child{F.G, F.gs, d1, F.M1, 0, nullptr, nullptr, nullptr, 0};
                J.stack.push_back(child);
                continue;
            }
            case 1:  // GG = coefficients d1..d of M1 * G[0..d]   (src/sigma_basis.rs:205-209)
                F.stage = 2;
                op.kind = 1;
                op.mul = MulDesc{F.M1, F.G, F.GG, F.gs, (uint32_t)(d1 + 1), (uint32_t)(F.d + 1), 0,
                                 (uint32_t)d1, (uint32_t)(F.d + 1), J.m, J.m, J.b};
                op.keep = F.SP ? &F : nullptr;
                op.given = nullptr;
                op.need = F.G == J.G ? F.d + 1 : 0;
                return true;
            case 2: {
                F.stage = 3;
                if (J.stack.size() >= J.stack.capacity()) {
                    set_error("PM-basis recursion deeper than %zu levels", J.stack.capacity());
                    err = RR_EINVAL;
                    return false;
                }
                Frame child{F.GG, d2 + 1, d2, F.M2, 0, nullptr, nullptr, nullptr, 0};
                J.stack.push_back(child);
                continue;
            }
            case 3:  // M = M2 * M1   (src/sigma_basis.rs:213)
                F.stage = 4;
                op.kind = 1;
                op.mul = MulDesc{F.M2, F.M1, F.M, d1 + 1, (uint32_t)(d2 + 1), (uint32_t)(d1 + 1), 0, 0,
                                 (uint32_t)(F.d + 1), J.m, J.m, J.m};
                op.keep = nullptr;
                op.given = F.sp_logn >= 0 ? &F : nullptr;
                op.need = 0;
                return true;
            default:
                J.pool.top = F.mark;
                J.stack.pop_back();
                continue;
        }
    }
    J.done = true;
    return false;
}

static uint32_t ntt_logn_for(const MulDesc &D) {
    uint32_t need = std::max<uint32_t>(D.hi, D.la + D.lb - 1 - D.lo), logn = 1;
    while ((1u << logn) < need) logn++;
    return logn;
}

// Direct (schoolbook) product or NTT product.  A direct product costs about
// la lb multiply-adds per output entry against 2 max(la, lb) for the NTT's
// pointwise part, but it is one launch instead of three: a round with few
// products (a PM-basis running alone, latency-bound) takes the direct form
// up to la lb = 256, a round with many (batched, throughput-bound) only up
// to 16.
// Mod any other prime a transform product runs three times (one per CRT
// prime) plus a reduction and a Garner pass, so the direct form goes four
// times as far.
constexpr size_t SCHOOL_MANY = 8;
static bool use_school(const MulDesc &D, uint32_t p, size_t round_products = 1) {
    uint64_t cost = (uint64_t)std::min(D.la, D.hi) * std::min(D.lb, D.hi);
    uint64_t lim = round_products >= SCHOOL_MANY ? 16u : 256u;
    if (p != NTT_PRIME) lim *= 4;
    return cost <= lim;
}

// Work space of one transform product: operand and product spectra; mod
// another prime also the operands reduced mod a smaller CRT prime and the
// three residue windows.
static size_t product_words(const MulDesc &D, uint32_t logn, uint32_t p) {
    const size_t N = (size_t)1 << logn, a64 = 64;
    size_t w = ((size_t)D.rows * D.inner * N + a64) + ((size_t)D.inner * D.cols * N + a64) +
               ((size_t)D.rows * D.cols * N + a64);
    if (p == NTT_PRIME) return w;
    const size_t la = std::min<size_t>(D.la, N), lb = std::min<size_t>(D.lb, N);
    w += ((size_t)D.rows * D.inner * la + a64) + ((size_t)D.inner * D.cols * lb + a64);
    w += 3 * ((size_t)D.rows * D.cols * (D.hi - D.lo) + a64);
    return w;
}

rr_status PmScheduler::ensure_scratch(size_t words) {
    if (words <= scratch_words) return RR_OK;
    RR_HIP(hipStreamSynchronize(st));
    scratch.~DevBuf();
    new (&scratch) DevBuf();
    size_t w = std::max(words, scratch_words * 2);
    rr_status s = scratch.alloc(w * 4);
    if (s != RR_OK) return s;
    scratch_words = w;
    return RR_OK;
}

rr_status PmScheduler::init(uint32_t p_, hipStream_t st_, uint32_t max_logn, uint32_t max_m, uint32_t max_n) {
    p = p_;
    st = st_;
    rr_status s;
    if ((s = init_tables(std::max<uint32_t>(2, std::min(max_logn, NTT_MAX_LOGN)))) != RR_OK) return s;
    // any mix of block widths: the largest leaf (b = 32: m = 64, K = 4) needs 150 KB
    lds = 0;
    for (uint32_t mm = 2; mm <= 128; mm += 2) {
        const uint32_t K = leaf_order(mm);
        lds = std::max(lds, leaf_lds_words(mm, mm / 2, K) * 4);
        RR_CHECK(K == 1 || (adv_items(mm, mm, K) <= ADV_ITEMS * (LEAF_THREADS / 64) &&
                            adv_items(mm, mm / 2, K) <= ADV_ITEMS * (LEAF_THREADS / 64)),
                 RR_EINVAL, "leaf of %u rows too large for the advance tiles", mm);
    }
    (void)max_m;
    (void)max_n;
    RR_CHECK(lds <= 160 * 1024, RR_EINVAL, "leaf needs %zu bytes of LDS", lds);
    RR_HIP(hipFuncSetAttribute((const void *)k_mbasis_iter, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    return RR_OK;
}

rr_status PmScheduler::init_tables(uint32_t logn) {
    rr_status s;
    if (p == NTT_PRIME) return tables.init(logn, st);
    if ((s = tables.init(logn, st, CRT_PRIMES[0], CRT_ROOTS[0])) != RR_OK) return s;
    if ((s = crt_tabs[0].init(logn, st, CRT_PRIMES[1], CRT_ROOTS[1])) != RR_OK) return s;
    return crt_tabs[1].init(logn, st, CRT_PRIMES[2], CRT_ROOTS[2]);
}

rr_status PmScheduler::add(PmJob *J) {
    J->m = 2 * J->b;
    if (!J->K) J->K = leaf_order(J->m);
    RR_CHECK(J->K >= 1 && J->K <= leaf_order(J->m), RR_EINVAL, "leaf order %u above %u", J->K, leaf_order(J->m));
    rr_status s;
    if ((s = J->pool.init(rec_words(J->d, J->K, J->m, J->b) + 1024)) != RR_OK) return s;
    RR_HIP(hipMemsetAsync(J->delta, 0, (size_t)J->m * 8, st));
    J->stack.clear();
    J->stack.reserve(PM_MAX_DEPTH);  // frames hold pointers to each other's spectra: no reallocation
    J->stack.push_back(Frame{J->G, J->d + 1, J->d, J->M, 0, nullptr, nullptr, nullptr, 0});
    J->done = false;
    J->has_op = false;
    J->err = RR_OK;
    return RR_OK;
}

// One round: the next operation of every unfinished job, launched together.
// A job whose next operation reads coefficients of G beyond those written so
// far (its sequence still being produced) waits.
rr_status PmScheduler::round(std::vector<PmJob *> &jobs, bool *launched) {
    const ModP mp = ModP::make(p);
    std::vector<LeafDesc> leaves;
    std::vector<MulDesc> school;
    std::vector<NttProduct> ntt;
    *launched = false;
    last.clear();
    // Products go first and leaves wait for a round in which no job has a
    // product pending, so that the leaves of many jobs share one launch (a
    // job reaches a leaf after at most its recursion depth of products).
    bool any_mul = false;
    waiting = 0;
    for (auto *J : jobs) {
        if (J->done) continue;
        if (!J->has_op) {
            Op op;
            rr_status e;
            if (!next_op(*J, op, e)) {
                if (e != RR_OK) {  // this job's own failure: it ends, the others go on
                    J->err = e;
                    J->done = true;
                }
                continue;
            }
            J->has_op = true;
            J->op_kind = op.kind;
            J->op_leaf = op.leaf;
            J->op_mul = op.mul;
            J->op_keep = op.keep;
            J->op_given = op.given;
            J->op_need = op.need;
        }
        if (J->op_need > J->avail) {
            waiting++;
            continue;
        }
        any_mul |= J->op_kind == 1;
    }
    auto ready = [](const PmJob *J) { return !J->done && J->has_op && J->op_need <= J->avail; };
    for (auto *J : jobs)
        RR_CHECK(!fail_tag || J->tag != fail_tag || !ready(J), RR_ENOMEM, "injected PM-basis round failure (job %llu)",
                 (unsigned long long)J->tag);
    size_t nmul = 0;
    if (any_mul)
        for (auto *J : jobs) nmul += ready(J) && J->op_kind == 1;
    if (any_mul) {  // the round's transform work space, before any job's operation is taken
        size_t words = 0;
        for (auto *J : jobs)
            if (ready(J) && J->op_kind == 1 && !use_school(J->op_mul, p, nmul))
                words += product_words(J->op_mul, ntt_logn_for(J->op_mul), p);
        rr_status e;
        if (words && (e = ensure_scratch(words)) != RR_OK) return e;
    }
    for (auto *J : jobs) {
        if (!ready(J) || J->op_kind != (any_mul ? 1 : 0)) continue;
        last.push_back(J);
        J->has_op = false;
        if (J->op_kind == 0) {
            leaves.push_back(J->op_leaf);
        } else if (use_school(J->op_mul, p, nmul)) {
            school.push_back(J->op_mul);
            if (J->op_keep) J->op_keep->sp_logn = -1;  // nothing kept: the later product transforms M1 itself
        } else {
            NttProduct o{ntt_logn_for(J->op_mul), J->op_mul, nullptr, nullptr};
            if (p != NTT_PRIME) {  // spectra mod three primes: none kept
                if (J->op_keep) J->op_keep->sp_logn = -1;
            } else if (J->op_keep && o.logn == node_logn(J->op_keep->d)) {
                o.keep = J->op_keep->SP;
                J->op_keep->sp_logn = (int)o.logn;
            } else if (J->op_keep) {
                J->op_keep->sp_logn = -1;
            }
            if (p == NTT_PRIME && J->op_given && J->op_given->sp_logn == (int)o.logn) o.given = J->op_given->SP;
            ntt.push_back(o);
        }
        J->op_keep = J->op_given = nullptr;
    }
    if (leaves.empty() && school.empty() && ntt.empty()) return RR_OK;
    *launched = true;
    rounds++;
    // one launch per LEAF_PACK leaves
    for (size_t i = 0; i < leaves.size();) {
        LeafPack pk;
        uint32_t cnt = 0;
        size_t lw = 0;
        while (i < leaves.size() && cnt < LEAF_PACK) {
            pk.d[cnt++] = leaves[i];
            lw = std::max(lw, leaf_lds_words(leaves[i].m, leaves[i].n, leaves[i].d));
            i++;
        }
        hipLaunchKernelGGL(k_mbasis_iter, dim3(cnt), dim3(LEAF_THREADS), lw * 4, st, pk, p, mp.pd, mp.inv);
        RR_HIP(hipGetLastError());
        launches++;
    }
    if (!school.empty()) {
        rr_status e = launch_school(school, p, st, &launches);
        if (e != RR_OK) return e;
    }
    if (!ntt.empty()) return run_ntt_products(ntt);
    return RR_OK;
}

rr_status PmScheduler::run(std::vector<PmJob *> &jobs) {
    uint32_t max_logn = 2, max_m = 0, max_n = 0;
    for (auto *J : jobs) {
        uint32_t l = 1;
        while ((1ull << l) < J->d + 1) l++;
        max_logn = std::max(max_logn, l);
        max_m = std::max(max_m, 2 * J->b);
        max_n = std::max(max_n, J->b);
    }
    rr_status s;
    if ((s = init(p, st, max_logn, max_m, max_n)) != RR_OK) return s;
    if (jobs.size() == 1 && !jobs[0]->K) jobs[0]->K = leaf_order_alone(2 * jobs[0]->b);
    for (auto *J : jobs)
        if ((s = add(J)) != RR_OK) return s;
    // RR_PROGRESS=<seconds>: a line on stderr at most that often (long calls)
    const char *pe = getenv("RR_PROGRESS");
    const double every = pe ? atof(pe) : 0.0;
    auto t_last = std::chrono::steady_clock::now();
    for (;;) {
        bool launched;
        if ((s = round(jobs, &launched)) != RR_OK) return s;
        if (!launched) break;
        if (every > 0) {
            const auto now = std::chrono::steady_clock::now();
            if (std::chrono::duration<double>(now - t_last).count() >= every) {
                t_last = now;
                fprintf(stderr, "rr_pm_basis: %llu rounds, %llu launches queued\n", (unsigned long long)rounds,
                        (unsigned long long)launches);
                fflush(stderr);
            }
        }
    }
    for (auto *J : jobs)
        if (J->err != RR_OK) return J->err;
    return RR_OK;
}

// Each NTT product: forward transforms of both operands into round scratch,
// pointwise products (with 1/N), inverse transforms into the output window.
rr_status PmScheduler::run_ntt_products(std::vector<NttProduct> &ops) {
    if (p != NTT_PRIME) return run_crt_products(ops);
    size_t words = 0;
    for (auto &o : ops) words += product_words(o.D, o.logn, p);
    rr_status s;
    if ((s = ensure_scratch(words)) != RR_OK) return s;
    // same-size products share their launches (forward, pointwise, inverse)
    std::stable_sort(ops.begin(), ops.end(),
                     [](const NttProduct &x, const NttProduct &y) { return x.logn < y.logn; });
    size_t off = 0;
    for (size_t g0 = 0; g0 < ops.size();) {
        const uint32_t logn = ops[g0].logn, N = 1u << logn;
        RR_CHECK(logn <= tables.logn, RR_EINVAL, "transform 2^%u above table 2^%u", logn, tables.logn);
        size_t g1 = g0;
        while (g1 < ops.size() && ops[g1].logn == logn && g1 - g0 < NTT_PACK / 2) g1++;
        NttFwdPack fp;
        PwPack pp;
        NttInvPack ip;
        uint32_t nf = 0, np = 0;
        for (size_t k = g0; k < g1; k++) {
            const MulDesc &D = ops[k].D;
            uint32_t *fa = ops[k].keep;
            if (!fa) {
                fa = scratch.as<uint32_t>() + off;
                off += ((size_t)D.rows * D.inner * N + 63) & ~size_t(63);
            }
            const uint32_t *fb = ops[k].given;
            uint32_t *fbw = nullptr;
            if (!fb) {
                fbw = scratch.as<uint32_t>() + off;
                off += ((size_t)D.inner * D.cols * N + 63) & ~size_t(63);
                fb = fbw;
            }
            uint32_t *fc = scratch.as<uint32_t>() + off;
            off += ((size_t)D.rows * D.cols * N + 63) & ~size_t(63);
            fp.d[nf++] = NttFwdDesc{D.A, D.la, 0, D.la, fa, D.rows * D.inner};
            if (fbw) fp.d[nf++] = NttFwdDesc{D.B, D.sb, D.ob, D.lb, fbw, D.inner * D.cols};
            pp.d[np] = PwDesc{fa, fb, fc, D.rows, D.inner, D.cols, h_invmod(N, NTT_PRIME)};
            ip.d[np] = NttInvDesc{fc, D.rows * D.cols, D.lo, D.hi - D.lo, D.C, D.hi - D.lo};
            np++;
        }
        if ((s = ntt_forward_pack(tables, fp, nf, logn, st)) != RR_OK) return s;
        if ((s = pointwise_pack(pp, np, N, st)) != RR_OK) return s;
        if ((s = ntt_inverse_pack(tables, ip, np, logn, st)) != RR_OK) return s;
        launches += 3;
        g0 = g1;
    }
    return RR_OK;
}

// Products mod a prime p other than the NTT prime: each product runs mod the
// three CRT primes (operands reduced first where p exceeds the prime: the
// butterflies need residues), and Garner's method takes the three residue
// windows to the exact coefficients, which are reduced mod p.  The exact
// coefficient in the window [lo, hi) sums at most inner min(la, lb, N)
// products below p^2 (wrapped terms fall below lo, operand coefficients at
// or beyond N only reach outputs at or beyond N >= hi), which must stay below
// the product of the primes (checked here; for example inner <= 128 with
// 2^21 terms of any p < 2^31, or any inner <= 128 and N <= 2^26 for p < 2^16).
rr_status PmScheduler::run_crt_products(std::vector<NttProduct> &ops) {
    const unsigned __int128 Q = (unsigned __int128)CRT_PRIMES[0] * CRT_PRIMES[1] * CRT_PRIMES[2];
    size_t words = 0;
    for (auto &o : ops) {
        const MulDesc &D = o.D;
        const uint64_t terms = std::min<uint64_t>(std::min(D.la, D.lb), 1ull << o.logn) * D.inner;
        RR_CHECK((unsigned __int128)terms * (p - 1) * (p - 1) < Q, RR_EINVAL,
                 "product of %u x %u entries exceeds the three-prime range", D.inner, std::min(D.la, D.lb));
        words += product_words(D, o.logn, p);
    }
    rr_status s;
    if ((s = ensure_scratch(words)) != RR_OK) return s;
    std::stable_sort(ops.begin(), ops.end(),
                     [](const NttProduct &x, const NttProduct &y) { return x.logn < y.logn; });
    uint32_t *base = scratch.as<uint32_t>();
    size_t off = 0;
    auto take = [&](size_t w) {
        uint32_t *r = base + off;
        off += (w + 63) & ~size_t(63);
        return r;
    };
    struct Win {
        uint32_t *fa, *fb, *fc, *ra, *rb, *res[3];
        uint32_t la, lb;
    };
    std::vector<Win> win(ops.size());
    for (size_t g0 = 0; g0 < ops.size();) {
        const uint32_t logn = ops[g0].logn, N = 1u << logn;
        RR_CHECK(logn <= tables.logn && logn <= crt_tabs[0].logn && logn <= crt_tabs[1].logn, RR_EINVAL,
                 "transform 2^%u above table 2^%u", logn, tables.logn);
        size_t g1 = g0;
        while (g1 < ops.size() && ops[g1].logn == logn && g1 - g0 < NTT_PACK / 2) g1++;
        for (size_t k = g0; k < g1; k++) {
            const MulDesc &D = ops[k].D;
            Win &W = win[k];
            W.la = std::min(D.la, N);
            W.lb = std::min(D.lb, N);
            W.fa = take((size_t)D.rows * D.inner * N);
            W.fb = take((size_t)D.inner * D.cols * N);
            W.fc = take((size_t)D.rows * D.cols * N);
            W.ra = take((size_t)D.rows * D.inner * W.la);
            W.rb = take((size_t)D.inner * D.cols * W.lb);
            for (auto &r : W.res) r = take((size_t)D.rows * D.cols * (D.hi - D.lo));
        }
        for (int t = 0; t < 3; t++) {
            const NttTables &T = t == 0 ? tables : crt_tabs[t - 1];
            const uint32_t q = T.prime;
            NttFwdPack fp;
            PwPack pp;
            NttInvPack ip;
            uint32_t nf = 0, np = 0;
            for (size_t k = g0; k < g1; k++) {
                const MulDesc &D = ops[k].D;
                const Win &W = win[k];
                if (p > q) {  // operands reduced mod q, compact
                    if ((s = mod_copy(D.A, D.la, 0, W.la, D.rows * D.inner, q, W.ra, st)) != RR_OK) return s;
                    if ((s = mod_copy(D.B, D.sb, D.ob, W.lb, D.inner * D.cols, q, W.rb, st)) != RR_OK) return s;
                    fp.d[nf++] = NttFwdDesc{W.ra, W.la, 0, W.la, W.fa, D.rows * D.inner};
                    fp.d[nf++] = NttFwdDesc{W.rb, W.lb, 0, W.lb, W.fb, D.inner * D.cols};
                    launches += 2;
                } else {
                    fp.d[nf++] = NttFwdDesc{D.A, D.la, 0, W.la, W.fa, D.rows * D.inner};
                    fp.d[nf++] = NttFwdDesc{D.B, D.sb, D.ob, W.lb, W.fb, D.inner * D.cols};
                }
                pp.d[np] = PwDesc{W.fa, W.fb, W.fc, D.rows, D.inner, D.cols, h_invmod(N % q, q)};
                ip.d[np] = NttInvDesc{W.fc, D.rows * D.cols, D.lo, D.hi - D.lo, W.res[t], D.hi - D.lo};
                np++;
            }
            if ((s = ntt_forward_pack(T, fp, nf, logn, st)) != RR_OK) return s;
            if ((s = pointwise_pack(pp, np, N, st, q)) != RR_OK) return s;
            if ((s = ntt_inverse_pack(T, ip, np, logn, st)) != RR_OK) return s;
            launches += 3;
        }
        for (size_t k = g0; k < g1; k++) {
            const MulDesc &D = ops[k].D;
            const Win &W = win[k];
            if ((s = garner3(W.res[0], W.res[1], W.res[2], (uint64_t)D.rows * D.cols * (D.hi - D.lo), p, D.C, st)) !=
                RR_OK)
                return s;
            launches++;
        }
        g0 = g1;
    }
    return RR_OK;
}

// Device PM-basis for several matrices at once.  For each job: G [m][b][d+1]
// on the device; writes M [m][m][d+1] (device) and delta [m] (device).
rr_status pm_basis_batch(uint32_t p, std::vector<PmJob *> &jobs, hipStream_t st, uint64_t *rounds,
                         uint64_t *launches) {
    PmScheduler S;
    S.p = p;
    S.st = st;
    rr_status s;
    if ((s = S.run(jobs)) != RR_OK) return s;
    if (rounds) *rounds = S.rounds;
    if (launches) *launches = S.launches;
    return RR_OK;
}

// Device PM-basis from a device-resident G [m][b][d+1]; writes M [m][m][d+1] (device) and delta (host).
rr_status pm_basis_device(uint32_t p, uint32_t b, uint64_t d, const uint32_t *G, uint32_t *M, int64_t *delta_host,
                          hipStream_t st) {
    PmJob J;
    J.b = b;
    J.d = d;
    J.G = G;
    J.M = M;
    DevBuf dl;
    rr_status s;
    if ((s = dl.alloc((size_t)2 * b * 8)) != RR_OK) return s;
    J.delta = dl.as<int64_t>();
    std::vector<PmJob *> jobs{&J};
    if ((s = pm_basis_batch(p, jobs, st, nullptr, nullptr)) != RR_OK) return s;
    RR_HIP(hipMemcpyAsync(delta_host, dl.ptr, (size_t)2 * b * 8, hipMemcpyDeviceToHost, st));
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

// G [2b][b][len] from the t-major upper-triangle sequence [L][P]
// (src/sigma_basis.rs:33-46): coefficients k0 <= k < k1 (coefficient k is
// S_k, which seq holds at entry k - s0)
__global__ void k_build_G(const uint32_t *seq, uint32_t b, uint64_t len, uint64_t k0, uint64_t k1, uint64_t s0,
                          uint32_t *G) {
    const uint32_t P = b * (b + 1) / 2;
    const uint64_t w = k1 - k0, total = (uint64_t)2 * b * b * w;
    for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k = k0 + idx % w, ij = idx / w;
        uint32_t j = (uint32_t)(ij % b), i = (uint32_t)(ij / b);
        uint32_t v;
        if (i < b) {
            uint32_t lo = min(i, j), hi = max(i, j);
            uint32_t ii = lo * b - lo * (lo - 1) / 2 + (hi - lo);
            v = seq[(k - s0) * P + ii];
        } else {
            v = (k == 0 && i - b == j) ? 1u : 0u;
        }
        G[ij * len + k] = v;
    }
}

rr_status build_G_range(const uint32_t *seq_dev, uint32_t b, uint64_t len, uint64_t k0, uint64_t k1, uint32_t *G,
                        hipStream_t st, uint64_t s0) {
    RR_CHECK(s0 <= k0 && k0 <= k1 && k1 <= len, RR_EINVAL, "coefficients [%llu, %llu) outside [%llu, %llu)",
             (unsigned long long)k0, (unsigned long long)k1, (unsigned long long)s0, (unsigned long long)len);
    if (k0 == k1) return RR_OK;
    const uint64_t total = (uint64_t)2 * b * b * (k1 - k0);
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((total + 255) / 256, 16384);
    hipLaunchKernelGGL(k_build_G, dim3(std::max(1u, blocks)), dim3(256), 0, st, seq_dev, b, len, k0, k1, s0, G);
    RR_HIP(hipGetLastError());
    return RR_OK;
}

rr_status build_G_device(const uint32_t *seq_dev, uint32_t b, uint64_t len, uint32_t *G, hipStream_t st) {
    return build_G_range(seq_dev, b, len, 0, len, G, st);
}

}  // namespace rr

using namespace rr;

extern "C" {

rr_status rr_pm_basis(const uint32_t *seq, uint32_t b, uint64_t L, uint64_t d, uint32_t p, uint32_t *basis_out,
                      int64_t *delta_out) {
    return rr::guarded([&]() -> rr_status {
        clear_error();
        RR_CHECK(seq && basis_out && delta_out, RR_EINVAL, "null argument");
        RR_CHECK(b >= 1 && b <= 64, RR_EINVAL, "b=%u must be in 1..64", b);
        RR_CHECK(p > 2 && p < (1u << 31) && h_is_prime(p), RR_EINVAL, "p=%u must be a prime in (2, 2^31)", p);
        RR_CHECK(d >= 1 && d < L, RR_EINVAL, "order d=%llu must satisfy 1 <= d < L=%llu", (unsigned long long)d,
                 (unsigned long long)L);
        RR_CHECK(d + 1 <= (1ull << NTT_MAX_LOGN), RR_EINVAL, "order d=%llu above the 2^%u-point transform limit",
                 (unsigned long long)d, NTT_MAX_LOGN);
        const uint32_t P = b * (b + 1) / 2, m = 2 * b;
        const uint64_t len = d + 1;
        hipStream_t st;
        RR_HIP(rr::stream_create(&st));
        struct S {
            hipStream_t s;
            ~S() { rr::stream_destroy(s); }
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
    });
}

rr_status rr_rank_from_basis(const uint32_t *basis, const int64_t *delta, uint32_t b, uint64_t len, uint32_t p,
                             int64_t *rank_out) {
    return rr::guarded([&]() -> rr_status {
        clear_error();
        RR_CHECK(basis && delta && rank_out, RR_EINVAL, "null argument");
        RR_CHECK(b >= 1 && b <= 64 && len >= 1, RR_EINVAL, "bad sizes");
        rr_status s;
        *rank_out = rank_from_basis_host(basis, delta, b, len, p, &s);
        return s;
    });
}

}  // extern "C"

extern "C" {

// Natural-order transforms in the reference convention (src/ntt.rs:164-211):
// data [count][2^logn], forward with root 31^((p-1)/n), inverse scaled by 1/n.
// The device transform is DIF (bit-reversed spectrum); the permutation to
// natural order is done on the host here, products never need it.
rr_status rr_ntt(uint32_t *data, uint32_t logn, uint32_t count, int inverse) {
    return rr::guarded([&]() -> rr_status {
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
        RR_HIP(rr::stream_create(&st));
        struct S {
            hipStream_t s;
            ~S() { rr::stream_destroy(s); }
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
    });
}

// Truncated polynomial-matrix product (src/poly_mat_mul.rs:89-256):
// C [rows][cols][hi-lo] = coefficients [lo, hi) of A [rows][inner][la] * B [inner][cols][lb], exact mod p.
rr_status rr_poly_mat_mul(const uint32_t *A, uint64_t la, const uint32_t *B, uint64_t lb, uint32_t rows,
                          uint32_t inner, uint32_t cols, uint64_t lo, uint64_t hi, uint32_t p, uint32_t *C) {
    return rr::guarded([&]() -> rr_status {
        clear_error();
        RR_CHECK(A && B && C && la >= 1 && lb >= 1 && hi > lo && rows && inner && cols, RR_EINVAL,
                 "bad product arguments");
        RR_CHECK(inner <= 128 && hi <= la + lb - 1 + 0ull, RR_EINVAL, "inner <= 128 and hi <= la + lb - 1");
        RR_CHECK(p > 2 && p < (1u << 31) && h_is_prime(p), RR_EINVAL, "p=%u must be a prime in (2, 2^31)", p);
        hipStream_t st;
        RR_HIP(rr::stream_create(&st));
        struct S {
            hipStream_t s;
            ~S() { rr::stream_destroy(s); }
        } guard{st};
        DevBuf dA, dB, dC;
        rr_status s;
        size_t sa = (size_t)rows * inner * la, sbn = (size_t)inner * cols * lb, sc = (size_t)rows * cols * (hi - lo);
        if ((s = dA.alloc(sa * 4)) != RR_OK) return s;
        if ((s = dB.alloc(sbn * 4)) != RR_OK) return s;
        if ((s = dC.alloc(sc * 4)) != RR_OK) return s;
        for (size_t i = 0; i < sa; i++) RR_CHECK(A[i] < p, RR_EINVAL, "A entry %zu >= p", i);
        for (size_t i = 0; i < sbn; i++) RR_CHECK(B[i] < p, RR_EINVAL, "B entry %zu >= p", i);
        RR_HIP(hipMemcpyAsync(dA.ptr, A, sa * 4, hipMemcpyHostToDevice, st));
        RR_HIP(hipMemcpyAsync(dB.ptr, B, sbn * 4, hipMemcpyHostToDevice, st));
        MulDesc D{dA.as<uint32_t>(), dB.as<uint32_t>(), dC.as<uint32_t>(), lb, (uint32_t)la, (uint32_t)lb, 0,
                  (uint32_t)lo, (uint32_t)hi, rows, inner, cols};
        PmScheduler S;
        S.p = p;
        S.st = st;
        if (use_school(D, p)) {
            if ((s = launch_school(std::vector<MulDesc>{D}, p, st, nullptr)) != RR_OK) return s;
        } else {
            uint32_t logn = ntt_logn_for(D);
            RR_CHECK(logn <= NTT_MAX_LOGN, RR_EINVAL, "product too long");
            if ((s = S.init_tables(logn)) != RR_OK) return s;
            std::vector<NttProduct> ops{{logn, D, nullptr, nullptr}};
            if ((s = S.run_ntt_products(ops)) != RR_OK) return s;
        }
        RR_HIP(hipMemcpyAsync(C, dC.ptr, sc * 4, hipMemcpyDeviceToHost, st));
        RR_HIP(hipStreamSynchronize(st));
        return RR_OK;
    });
}

}  // extern "C"
