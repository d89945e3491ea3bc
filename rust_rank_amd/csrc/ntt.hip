// ntt.hip -- number-theoretic transforms and polynomial-matrix products mod
// p = 2013265921 (15 * 2^27 + 1, generator 31) on gfx950.
//
// Replaces the reference's ntt / matrix_ntt_parallel (src/ntt.rs:164-269) and
// poly_mat_mul_fft_red (src/poly_mat_mul.rs:113-256); the CUDA twin is
// cuda_test/cudantt.h:110-415 + cupoly_mat_mul.h:138-305.  The reference lifts
// its operands to a 61-bit prime and reduces afterwards, which is only exact
// for small p; here the transforms run directly mod p, so every product is
// exact.
//
// Design: one workgroup transforms whole polynomials in LDS (up to 2^14 words
// = 64 KB per workgroup; several short polynomials share a workgroup).  The
// forward transform is decimation-in-frequency (natural order in, bit-reversed
// out), the inverse is decimation-in-time (bit-reversed in, natural out), so a
// product never permutes.  Twiddles come from a table w^k, k < N/2, with
// Shoup companions floor(w^k 2^32 / p), so a twiddle multiply is one mulhi and
// two mullo.  The 1/N of the inverse is folded into the pointwise product.
#include "common.hpp"
#include "ntt.hpp"
#include <algorithm>
#include <mutex>

namespace rr {

constexpr int NTT_THREADS = 256;
constexpr uint32_t NTT_WG_WORDS = 2048;    // shorter transforms share a workgroup up to this many words
constexpr uint32_t NTT_U = 4;              // butterflies per thread in flight

__host__ __device__ constexpr uint32_t ntt_per_wg(uint32_t N) { return N >= NTT_WG_WORDS ? 1 : NTT_WG_WORDS / N; }

// One radix-2 stage over `total` butterflies of span h = 2^lh in LDS.  Every
// thread issues NTT_U butterflies' LDS and twiddle loads before it computes;
// butterflies past `total` repeat the thread's first one and are not stored.
template <bool DIF>
__device__ __forceinline__ void lds_stage(uint32_t *s, uint32_t total, uint32_t logn, uint32_t lh, const uint32_t *tw,
                                          const uint32_t *twq, uint32_t tstride, uint32_t p, uint32_t ps, uint32_t tsh) {
    const uint32_t h = 1u << lh, N = 1u << logn, half = N >> 1;
    for (uint32_t t0 = threadIdx.x; t0 < total; t0 += NTT_THREADS * NTT_U) {
        uint32_t i0[NTT_U], x0[NTT_U], x1[NTT_U], w[NTT_U], wq[NTT_U];
#pragma unroll
        for (uint32_t u = 0; u < NTT_U; u++) {
            uint32_t t = t0 + u * NTT_THREADS;
            t = t < total ? t : t0;
            const uint32_t poly = t >> (logn - 1), bfl = t & (half - 1);
            const uint32_t j = bfl & (h - 1), blk = bfl >> lh;
            i0[u] = poly * ps + blk * 2 * h + j;
            x0[u] = s[i0[u]];
            x1[u] = s[i0[u] + h];
            const uint32_t e = j * tstride;
            w[u] = tw[e + (e >> tsh)];
            wq[u] = twq[e + (e >> tsh)];
        }
#pragma unroll
        for (uint32_t u = 0; u < NTT_U; u++) {
            uint32_t y0, y1;
            if constexpr (DIF) {
                y0 = addmod(x0[u], x1[u], p);
                y1 = shoup_mul(submod(x0[u], x1[u], p), w[u], wq[u], p);
            } else {
                const uint32_t bt = shoup_mul(x1[u], w[u], wq[u], p);
                y0 = addmod(x0[u], bt, p);
                y1 = submod(x0[u], bt, p);
            }
            if (t0 + u * NTT_THREADS < total) {
                s[i0[u]] = y0;
                s[i0[u] + h] = y1;
            }
        }
    }
    __syncthreads();
}

// Two radix-2 stages at once (spans 2g and g), one radix-4 butterfly per
// thread and step: the four words base + j + k g, k < 4, of a block of 4g
// (j < g).  DIF: the span-2g pairs (0, 2), (1, 3) with w_4g^j, w_4g^(j+g),
// then the span-g pairs (0, 1), (2, 3) with w_2g^j; DIT the same in reverse
// (span g first).  Every word sees the operations of the two radix-2 stages
// in the same order, with half the LDS passes and barriers.
template <bool DIF>
__device__ __forceinline__ void lds_stage4(uint32_t *s, uint32_t total4, uint32_t logn, uint32_t lg,
                                           const uint32_t *tw, const uint32_t *twq, uint32_t tstride4, uint32_t p,
                                           uint32_t ps, uint32_t tsh) {
    constexpr uint32_t U4 = 2;  // radix-4 butterflies per thread in flight
    const uint32_t g = 1u << lg, quarter = 1u << (logn - 2);
    for (uint32_t t0 = threadIdx.x; t0 < total4; t0 += NTT_THREADS * U4) {
        uint32_t i0[U4], x[U4][4], w[U4][3], wq[U4][3];
#pragma unroll
        for (uint32_t u = 0; u < U4; u++) {
            uint32_t t = t0 + u * NTT_THREADS;
            t = t < total4 ? t : t0;
            const uint32_t poly = t >> (logn - 2), bfl = t & (quarter - 1);
            const uint32_t j = bfl & (g - 1), blk = bfl >> lg;
            i0[u] = poly * ps + blk * 4 * g + j;
#pragma unroll
            for (uint32_t k = 0; k < 4; k++) x[u][k] = s[i0[u] + k * g];
            // w_4g^j, w_4g^(j+g), w_2g^j = w_4g^(2j)
            const uint32_t e[3] = {j * tstride4, (j + g) * tstride4, 2 * j * tstride4};
#pragma unroll
            for (uint32_t k = 0; k < 3; k++) {
                w[u][k] = tw[e[k] + (e[k] >> tsh)];
                wq[u][k] = twq[e[k] + (e[k] >> tsh)];
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < U4; u++) {
            uint32_t y[4];
            if constexpr (DIF) {
                const uint32_t b0 = addmod(x[u][0], x[u][2], p);
                const uint32_t b2 = shoup_mul(submod(x[u][0], x[u][2], p), w[u][0], wq[u][0], p);
                const uint32_t b1 = addmod(x[u][1], x[u][3], p);
                const uint32_t b3 = shoup_mul(submod(x[u][1], x[u][3], p), w[u][1], wq[u][1], p);
                y[0] = addmod(b0, b1, p);
                y[1] = shoup_mul(submod(b0, b1, p), w[u][2], wq[u][2], p);
                y[2] = addmod(b2, b3, p);
                y[3] = shoup_mul(submod(b2, b3, p), w[u][2], wq[u][2], p);
            } else {
                const uint32_t t1 = shoup_mul(x[u][1], w[u][2], wq[u][2], p);
                const uint32_t b0 = addmod(x[u][0], t1, p), b1 = submod(x[u][0], t1, p);
                const uint32_t t3 = shoup_mul(x[u][3], w[u][2], wq[u][2], p);
                const uint32_t b2 = addmod(x[u][2], t3, p), b3 = submod(x[u][2], t3, p);
                const uint32_t t2 = shoup_mul(b2, w[u][0], wq[u][0], p);
                y[0] = addmod(b0, t2, p);
                y[2] = submod(b0, t2, p);
                const uint32_t t4 = shoup_mul(b3, w[u][1], wq[u][1], p);
                y[1] = addmod(b1, t4, p);
                y[3] = submod(b1, t4, p);
            }
            if (t0 + u * NTT_THREADS < total4) {
#pragma unroll
                for (uint32_t k = 0; k < 4; k++) s[i0[u] + k * g] = y[k];
            }
        }
    }
    __syncthreads();
}

// In-LDS DIF over `npoly` polynomials of length N = 2^logn, ps words apart
// (N, or N + 1 where a transposed fill must avoid bank conflicts): radix-4
// stages from the largest spans down, one radix-2 stage of span 1 when logn
// is odd.  Twiddle k of the table is word k + (k >> tsh): tsh = TW_PAD for
// the tables staged in LDS (one spare word per 64, so the power-of-two
// strides of the middle stages spread over the banks), 31 for global ones.
__device__ void lds_dif(uint32_t *s, uint32_t npoly, uint32_t logn, const uint32_t *tw, const uint32_t *twq,
                        uint32_t tw_logn, uint32_t p, uint32_t ps, uint32_t tsh) {
    const uint32_t total = npoly << (logn - 1);
    uint32_t lh = logn;  // spans 2^(lh-1) and 2^(lh-2) next; w_{2h}^j = w_Ntab^(j * Ntab / 2h)
    for (; lh >= 2; lh -= 2)
        lds_stage4<true>(s, npoly << (logn - 2), logn, lh - 2, tw, twq, 1u << (tw_logn - lh), p, ps, tsh);
    if (lh == 1) lds_stage<true>(s, total, logn, 0, tw, twq, 1u << (tw_logn - 1), p, ps, tsh);
}

// In-LDS DIT (inverse twiddles in tw), bit-reversed in, natural out, unscaled:
// radix-4 stages from span 1 up, one radix-2 stage of span N/2 when logn is odd.
__device__ void lds_dit(uint32_t *s, uint32_t npoly, uint32_t logn, const uint32_t *tw, const uint32_t *twq,
                        uint32_t tw_logn, uint32_t p, uint32_t ps, uint32_t tsh) {
    const uint32_t total = npoly << (logn - 1);
    uint32_t lh = 0;  // spans 2^lh and 2^(lh+1) next
    for (; lh + 2 <= logn; lh += 2)
        lds_stage4<false>(s, npoly << (logn - 2), logn, lh, tw, twq, 1u << (tw_logn - lh - 2), p, ps, tsh);
    if (lh < logn) lds_stage<false>(s, total, logn, lh, tw, twq, 1u << (tw_logn - lh - 1), p, ps, tsh);
}

constexpr uint32_t TW_PAD = 6;  // one spare word per 64 in LDS twiddle tables
constexpr uint32_t TW_GLOBAL = 31;
__host__ __device__ constexpr uint32_t tw_lds_words(uint32_t n) { return n + (n >> TW_PAD); }

constexpr uint32_t NTT_LD = 8;  // global loads per thread in flight when staging

// Forward transforms: dst[q][0..N) = DIF(src[q][0..len) zero-padded), for the
// polynomials q of a [rows][cols] matrix with source stride `sstride` (words
// between consecutive polynomials) and source offset `soff` (first
// coefficient).  blockIdx.y selects the descriptor of the pack.
__global__ __launch_bounds__(NTT_THREADS) void k_ntt_fwd(NttFwdPack pk, uint32_t logn, const uint32_t *tw,
                                                         const uint32_t *twq, uint32_t tw_logn, uint32_t p) {
    extern __shared__ uint32_t s[];
    const NttFwdDesc &D = pk.d[blockIdx.y];
    const uint32_t N = 1u << logn, per = ntt_per_wg(N);
    const uint32_t q0 = blockIdx.x * per;
    if (q0 >= D.count) return;
    const uint32_t np = min(per, D.count - q0), words = np * N;
    const uint32_t *src = D.src + (size_t)q0 * D.sstride + D.soff;
    // lanes past the data or the source length load the first word and store zero
    for (uint32_t i0 = threadIdx.x; i0 < words; i0 += NTT_THREADS * NTT_LD) {
        uint32_t v[NTT_LD];
#pragma unroll
        for (uint32_t e = 0; e < NTT_LD; e++) {
            const uint32_t i = i0 + e * NTT_THREADS, q = i >> logn, k = i & (N - 1);
            const bool ok = i < words && k < D.len;
            v[e] = D.len ? src[ok ? (size_t)q * D.sstride + k : 0] : 0u;
        }
#pragma unroll
        for (uint32_t e = 0; e < NTT_LD; e++) {
            const uint32_t i = i0 + e * NTT_THREADS;
            if (i < words) s[i] = (i & (N - 1)) < D.len ? v[e] : 0u;
        }
    }
    // the N-point twiddles w_N^k, k < N/2, into LDS behind the data
    uint32_t *tws = s + max(N, NTT_WG_WORDS), *twsq = tws + tw_lds_words(N >> 1);
    const uint32_t tstep = 1u << (tw_logn - logn);
    for (uint32_t k = threadIdx.x; k < (N >> 1); k += NTT_THREADS) {
        tws[k + (k >> TW_PAD)] = tw[k * tstep];
        twsq[k + (k >> TW_PAD)] = twq[k * tstep];
    }
    __syncthreads();
    lds_dif(s, np, logn, tws, twsq, logn, p, N, TW_PAD);
    uint32_t *dst = D.dst + (size_t)q0 * N;
    for (uint32_t i0 = threadIdx.x; i0 < words; i0 += NTT_THREADS * NTT_LD) {
        uint32_t v[NTT_LD];
#pragma unroll
        for (uint32_t e = 0; e < NTT_LD; e++) v[e] = s[min(i0 + e * NTT_THREADS, words - 1)];
#pragma unroll
        for (uint32_t e = 0; e < NTT_LD; e++)
            if (i0 + e * NTT_THREADS < words) dst[i0 + e * NTT_THREADS] = v[e];
    }
}

// Inverse transforms of `count` spectra (already scaled by 1/N), writing
// coefficients [lo, lo + outlen) of each to dst[q * dstride + k - lo].
__global__ __launch_bounds__(NTT_THREADS) void k_ntt_inv(NttInvPack pk, uint32_t logn, const uint32_t *tw,
                                                         const uint32_t *twq, uint32_t tw_logn, uint32_t p) {
    extern __shared__ uint32_t s[];
    const NttInvDesc &D = pk.d[blockIdx.y];
    const uint32_t N = 1u << logn, per = ntt_per_wg(N);
    const uint32_t q0 = blockIdx.x * per;
    if (q0 >= D.count) return;
    const uint32_t np = min(per, D.count - q0), words = np * N;
    const uint32_t *src = D.src + (size_t)q0 * N;
    for (uint32_t i0 = threadIdx.x; i0 < words; i0 += NTT_THREADS * NTT_LD) {
        uint32_t v[NTT_LD];
#pragma unroll
        for (uint32_t e = 0; e < NTT_LD; e++) v[e] = src[min(i0 + e * NTT_THREADS, words - 1)];
#pragma unroll
        for (uint32_t e = 0; e < NTT_LD; e++)
            if (i0 + e * NTT_THREADS < words) s[i0 + e * NTT_THREADS] = v[e];
    }
    uint32_t *tws = s + max(N, NTT_WG_WORDS), *twsq = tws + tw_lds_words(N >> 1);
    const uint32_t tstep = 1u << (tw_logn - logn);
    for (uint32_t k = threadIdx.x; k < (N >> 1); k += NTT_THREADS) {
        tws[k + (k >> TW_PAD)] = tw[k * tstep];
        twsq[k + (k >> TW_PAD)] = twq[k * tstep];
    }
    __syncthreads();
    lds_dit(s, np, logn, tws, twsq, logn, p, N, TW_PAD);
    const uint32_t outs = np * D.outlen;
    for (uint32_t i0 = threadIdx.x; i0 < outs; i0 += NTT_THREADS * NTT_LD) {
        uint32_t v[NTT_LD];
#pragma unroll
        for (uint32_t e = 0; e < NTT_LD; e++) {
            const uint32_t i = min(i0 + e * NTT_THREADS, outs - 1), q = i / D.outlen, k = i - q * D.outlen;
            v[e] = s[q * N + D.lo + k];
        }
#pragma unroll
        for (uint32_t e = 0; e < NTT_LD; e++) {
            const uint32_t i = i0 + e * NTT_THREADS, q = i / D.outlen, k = i - q * D.outlen;
            if (i < outs) D.dst[(size_t)(q0 + q) * D.dstride + k] = v[e];
        }
    }
}

// ---------------------------------------------------------------------------
// Transforms above 2^14 points (up to 2^26): four steps over N = N1 N2, with
// element n = n1 + N1 n2; N2 = 2^ceil(logn/2) and N1 = 2^floor(logn/2) up to
// 2^20 points, above that N1 = 2^14 (the rows stay one-workgroup transforms)
// and N2 = N / 2^14.
// Forward: N2-point DIF along n2 for each n1 (column kernel, a tile of TC
// columns transposed into LDS), the twiddle w_N^(n1 k2), then N1-point DIF
// along n1 for each k2 (rows, contiguous, the ordinary kernel in place).
// Spectrum entry N1 j + j1 holds X[brev(j) + N2 brev(j1)]; the inverse takes
// the same order back (rows, then the conjugate twiddle, then columns), so
// a product never permutes.
__host__ __device__ inline uint32_t big_l2(uint32_t logn) { return logn <= 20 ? (logn + 1) / 2 : logn - NTT_LDS_LOGN; }
__host__ __device__ inline uint32_t big_ltc(uint32_t logn) {  // log2 of the columns per workgroup
    const uint32_t l1 = logn - big_l2(logn), byl = 13 - big_l2(logn);  // (N2 + 1) TC words, about 32 KB
    return l1 < byl ? l1 : byl;
}
__host__ __device__ inline uint32_t big_tc(uint32_t logn) { return 1u << big_ltc(logn); }

// w_N^e * x for e < N, from the table w_T^k, k < T/2 (T = 2^tw_logn >= N);
// w_T^(k + T/2) = -w_T^k.
__device__ __forceinline__ uint32_t tw_mul(uint32_t x, uint32_t e, uint32_t logn, const uint32_t *tw,
                                           uint32_t tw_logn, uint32_t p, double inv) {
    const uint32_t k = e << (tw_logn - logn), half = 1u << (tw_logn - 1);
    const uint32_t w = tw[k & (half - 1)];
    const uint32_t r = mulmod(x, w, p, inv);
    return (k & half) && r ? p - r : r;
}

__global__ __launch_bounds__(NTT_THREADS) void k_ntt_cols_fwd(NttFwdPack pk, uint32_t logn, const uint32_t *tw,
                                                              const uint32_t *twq, uint32_t tw_logn, uint32_t p,
                                                              double inv) {
    extern __shared__ uint32_t s[];
    const NttFwdDesc &D = pk.d[blockIdx.y];
    const uint32_t l2 = big_l2(logn), l1 = logn - l2, N1 = 1u << l1, N2 = 1u << l2;
    const uint32_t ltc = big_ltc(logn), TC = 1u << ltc;
    const uint32_t tiles = N1 / TC, q = blockIdx.x / tiles, c0 = (blockIdx.x - q * tiles) * TC;
    if (q >= D.count) return;
    const uint32_t ps = N2 + 1, words = TC * N2;
    const uint32_t *src = D.src + (size_t)q * D.sstride + D.soff;
    for (uint32_t i0 = threadIdx.x; i0 < words; i0 += NTT_THREADS * NTT_LD) {
        uint32_t v[NTT_LD];
#pragma unroll
        for (uint32_t e = 0; e < NTT_LD; e++) {
            const uint32_t i = i0 + e * NTT_THREADS, c = i & (TC - 1), n = c0 + c + N1 * (i >> ltc);
            const bool ok = i < words && n < D.len;
            v[e] = D.len ? src[ok ? n : 0] : 0u;
        }
#pragma unroll
        for (uint32_t e = 0; e < NTT_LD; e++) {
            const uint32_t i = i0 + e * NTT_THREADS, c = i & (TC - 1), n2 = i >> ltc;
            if (i < words) s[c * ps + n2] = (c0 + c + N1 * n2) < D.len ? v[e] : 0u;
        }
    }
    __syncthreads();
    lds_dif(s, TC, l2, tw, twq, tw_logn, p, ps, TW_GLOBAL);
    uint32_t *dst = D.dst + ((size_t)q << logn);
    for (uint32_t i = threadIdx.x; i < words; i += NTT_THREADS) {
        const uint32_t c = i & (TC - 1), j = i >> ltc, n1 = c0 + c;
        const uint32_t k2 = __brev(j) >> (32 - l2);
        dst[n1 + N1 * j] = tw_mul(s[c * ps + j], n1 * k2, logn, tw, tw_logn, p, inv);
    }
}

__global__ __launch_bounds__(NTT_THREADS) void k_ntt_cols_inv(NttInvPack pk, uint32_t logn, const uint32_t *tw,
                                                              const uint32_t *twq, uint32_t tw_logn, uint32_t p,
                                                              double inv) {
    extern __shared__ uint32_t s[];
    const NttInvDesc &D = pk.d[blockIdx.y];
    const uint32_t l2 = big_l2(logn), l1 = logn - l2, N1 = 1u << l1, N2 = 1u << l2;
    const uint32_t ltc = big_ltc(logn), TC = 1u << ltc;
    const uint32_t tiles = N1 / TC, q = blockIdx.x / tiles, c0 = (blockIdx.x - q * tiles) * TC;
    if (q >= D.count) return;
    const uint32_t ps = N2 + 1, words = TC * N2;
    const uint32_t *src = D.src + ((size_t)q << logn);
    for (uint32_t i = threadIdx.x; i < words; i += NTT_THREADS) {
        const uint32_t c = i & (TC - 1), j = i >> ltc, n1 = c0 + c;
        const uint32_t k2 = __brev(j) >> (32 - l2);
        s[c * ps + j] = tw_mul(src[n1 + N1 * j], n1 * k2, logn, tw, tw_logn, p, inv);  // tw holds w^-k here
    }
    __syncthreads();
    lds_dit(s, TC, l2, tw, twq, tw_logn, p, ps, TW_GLOBAL);
    uint32_t *dst = D.dst + (size_t)q * D.dstride;
    for (uint32_t i = threadIdx.x; i < words; i += NTT_THREADS) {
        const uint32_t c = i & (TC - 1), n2 = i >> ltc, n = c0 + c + N1 * n2;
        if (n >= D.lo && n - D.lo < D.outlen) dst[n - D.lo] = s[c * ps + n2];
    }
}

// Pointwise products: C[i][j][l] = scale * sum_r A[i][r][l] B[r][j][l] over
// spectra of length N (a power of two).  A thread owns a TI x TJ tile of C at
// one point l, so each step over r loads TI + TJ words for TI TJ products;
// threads run over (point, row tile, column tile, descriptor), points
// fastest, so short spectra still fill every lane and a wavefront's loads are
// runs of consecutive words.  Sums are exact in fp64: A is split into 16-bit
// halves (products < 2^47), folded every 64 terms.  Rows and columns past a
// descriptor's matrix re-read its last row / column and are not stored.
template <int TI, int TJ>
__global__ __launch_bounds__(256) void k_pointwise(PwPack pk, uint32_t ndesc, uint32_t logn, uint32_t rt,
                                                   uint32_t ct, uint32_t p, double pd, double inv) {
    const uint32_t N = 1u << logn;
    const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const uint32_t l = (uint32_t)(t & (N - 1));
    uint64_t q = t >> logn;
    const uint32_t it = (uint32_t)(q % rt);
    q /= rt;
    const uint32_t jt = (uint32_t)(q % ct);
    q /= ct;
    if (q >= ndesc) return;
    const PwDesc &D = pk.d[q];
    const uint32_t rows = D.rows, inner = D.inner, cols = D.cols, i0 = it * TI, j0 = jt * TJ;
    if (i0 >= rows || j0 >= cols) return;
    const GLOBAL_AS uint32_t *ap[TI];
    const GLOBAL_AS uint32_t *bp[TJ];
#pragma unroll
    for (int a = 0; a < TI; a++) ap[a] = as_global(D.A) + (size_t)min(i0 + a, rows - 1) * inner * N + l;
#pragma unroll
    for (int c = 0; c < TJ; c++) bp[c] = as_global(D.B) + (size_t)min(j0 + c, cols - 1) * N + l;
    const size_t bstep = (size_t)cols * N;
    double hi[TI][TJ], lo[TI][TJ];
#pragma unroll
    for (int a = 0; a < TI; a++)
#pragma unroll
        for (int c = 0; c < TJ; c++) hi[a][c] = lo[a][c] = 0.0;
    for (uint32_t r0 = 0; r0 < inner; r0 += 64) {
        const uint32_t re = min(inner, r0 + 64);
        for (uint32_t r = r0; r < re; r++) {
            uint32_t av[TI], bv[TJ];
#pragma unroll
            for (int a = 0; a < TI; a++) av[a] = ap[a][(size_t)r * N];
#pragma unroll
            for (int c = 0; c < TJ; c++) bv[c] = bp[c][r * bstep];
#pragma unroll
            for (int a = 0; a < TI; a++) {
                const double ah = (double)(av[a] >> 16), al = (double)(av[a] & 0xFFFFu);
#pragma unroll
                for (int c = 0; c < TJ; c++) {
                    const double bd = (double)bv[c];
                    hi[a][c] = fma(ah, bd, hi[a][c]);
                    lo[a][c] = fma(al, bd, lo[a][c]);
                }
            }
        }
#pragma unroll
        for (int a = 0; a < TI; a++)
#pragma unroll
            for (int c = 0; c < TJ; c++) {
                hi[a][c] = fold(hi[a][c], pd, inv);
                lo[a][c] = fold(lo[a][c], pd, inv);
            }
    }
    GLOBAL_AS uint32_t *C = as_global(D.C);
#pragma unroll
    for (int a = 0; a < TI; a++)
#pragma unroll
        for (int c = 0; c < TJ; c++)
            if (i0 + a < rows && j0 + c < cols) {
                const uint32_t v = (uint32_t)fold(fma(hi[a][c], 65536.0, lo[a][c]), pd, inv);
                C[((size_t)(i0 + a) * cols + j0 + c) * N + l] = mulmod(v, D.scale, p, inv);
            }
}

// dst[i] = src[(i / len) * sstride + soff + i % len] mod q (residues < 2^32 in, < q out)
__global__ void k_mod_copy(const uint32_t *src, uint64_t sstride, uint32_t soff, uint32_t len, uint64_t n, uint32_t q,
                           double qd, double qinv, uint32_t *dst) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t poly = i / len, k = i - poly * len;
        dst[i] = (uint32_t)fold((double)src[poly * sstride + soff + k], qd, qinv);
    }
}

// Garner: x = v0 + q0 v1 + q0 q1 v2 with v0 = r0, v1 = (r1 - v0) q0^-1 mod q1,
// v2 = (r2 - v0 - q0 v1) (q0 q1)^-1 mod q2; then x mod p from the digits.
struct Garner3 {
    uint32_t q0, q1, q2, inv01, inv012, q0m2, q01p, q0p, p;
    double i1, i2, ip;
};
__global__ void k_garner3(const uint32_t *r0, const uint32_t *r1, const uint32_t *r2, uint64_t n, Garner3 g,
                          uint32_t *C) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t v0 = r0[i];
        const uint32_t v0m1 = v0 % g.q1, v0m2 = v0 % g.q2;
        const uint32_t v1 = mulmod(submod(r1[i], v0m1, g.q1), g.inv01, g.q1, g.i1);
        const uint32_t t = submod(submod(r2[i], v0m2, g.q2), mulmod(g.q0m2, v1 % g.q2, g.q2, g.i2), g.q2);
        const uint32_t v2 = mulmod(t, g.inv012, g.q2, g.i2);
        uint32_t x = v0 % g.p;
        x = addmod(x, mulmod(g.q0p, v1 % g.p, g.p, g.ip), g.p);
        x = addmod(x, mulmod(g.q01p, v2 % g.p, g.p, g.ip), g.p);
        C[i] = x;
    }
}

// ---------------------------------------------------------------------------
// host side

rr_status mod_copy(const uint32_t *src, uint64_t sstride, uint32_t soff, uint32_t len, uint32_t count, uint32_t q,
                   uint32_t *dst, hipStream_t st) {
    const uint64_t n = (uint64_t)len * count;
    if (!n) return RR_OK;
    const ModP mq = ModP::make(q);
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_mod_copy, dim3(grid), dim3(256), 0, st, src, sstride, soff, len, n, q, mq.pd, mq.inv, dst);
    RR_HIP(hipGetLastError());
    return RR_OK;
}

rr_status garner3(const uint32_t *r0, const uint32_t *r1, const uint32_t *r2, uint64_t n, uint32_t p, uint32_t *C,
                  hipStream_t st) {
    if (!n) return RR_OK;
    Garner3 g;
    g.q0 = CRT_PRIMES[0];
    g.q1 = CRT_PRIMES[1];
    g.q2 = CRT_PRIMES[2];
    g.p = p;
    g.inv01 = h_invmod(g.q0 % g.q1, g.q1);
    g.inv012 = h_invmod(h_mulmod(g.q0 % g.q2, g.q1 % g.q2, g.q2), g.q2);
    g.q0m2 = g.q0 % g.q2;
    g.q0p = g.q0 % p;
    g.q01p = h_mulmod(g.q0 % p, g.q1 % p, p);
    g.i1 = ModP::make(g.q1).inv;
    g.i2 = ModP::make(g.q2).inv;
    g.ip = ModP::make(p).inv;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_garner3, dim3(grid), dim3(256), 0, st, r0, r1, r2, n, g, C);
    RR_HIP(hipGetLastError());
    return RR_OK;
}

rr_status NttTables::init(uint32_t logn_max, hipStream_t stream, uint32_t prime_, uint32_t root_) {
    prime = prime_;
    root = root_;
    const uint32_t p = prime;
    RR_CHECK(((p - 1) >> logn_max) << logn_max == p - 1, RR_EINVAL, "2^%u does not divide %u - 1", logn_max, p);
    logn = logn_max;
    uint32_t half = 1u << (logn - 1);
    if (logn == 0) half = 1;
    std::vector<uint32_t> f(half), fq(half), iv(half), ivq(half);
    uint32_t w = h_powmod(root, (p - 1) >> logn, p), wi = h_invmod(w, p);
    uint64_t a = 1, b = 1;
    for (uint32_t k = 0; k < half; k++) {
        f[k] = (uint32_t)a;
        iv[k] = (uint32_t)b;
        fq[k] = (uint32_t)(((uint64_t)f[k] << 32) / p);
        ivq[k] = (uint32_t)(((uint64_t)iv[k] << 32) / p);
        a = a * w % p;
        b = b * wi % p;
    }
    rr_status s;
    if ((s = fwd.alloc(half * 4)) != RR_OK) return s;
    if ((s = fwdq.alloc(half * 4)) != RR_OK) return s;
    if ((s = inv.alloc(half * 4)) != RR_OK) return s;
    if ((s = invq.alloc(half * 4)) != RR_OK) return s;
    RR_HIP(hipMemcpyAsync(fwd.ptr, f.data(), half * 4, hipMemcpyHostToDevice, stream));
    RR_HIP(hipMemcpyAsync(fwdq.ptr, fq.data(), half * 4, hipMemcpyHostToDevice, stream));
    RR_HIP(hipMemcpyAsync(inv.ptr, iv.data(), half * 4, hipMemcpyHostToDevice, stream));
    RR_HIP(hipMemcpyAsync(invq.ptr, ivq.data(), half * 4, hipMemcpyHostToDevice, stream));
    RR_HIP(hipStreamSynchronize(stream));
    tf = fwd.as<uint32_t>();
    tfq = fwdq.as<uint32_t>();
    ti = inv.as<uint32_t>();
    tiq = invq.as<uint32_t>();
    return RR_OK;
}

static inline uint32_t blocks_for(uint32_t count, uint32_t logn) {
    const uint32_t per = ntt_per_wg(1u << logn);
    return (count + per - 1) / per;
}

// data (one polynomial, or several sharing NTT_WG_WORDS) and the N/2 twiddle pairs
static inline size_t lds_for(uint32_t logn) {
    return ((size_t)std::max(1u << logn, NTT_WG_WORDS) + 2 * tw_lds_words(1u << (logn - 1))) * 4;
}

// Dynamic LDS above 64 KB needs the kernels' attribute raised (once per process).
static rr_status ntt_lds_attrs() {
    static std::once_flag once;
    static hipError_t err = hipSuccess;
    std::call_once(once, [] {
        const int bytes = (int)lds_for(NTT_LDS_LOGN);
        err = hipFuncSetAttribute((const void *)k_ntt_fwd, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        if (err == hipSuccess)
            err = hipFuncSetAttribute((const void *)k_ntt_inv, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    });
    RR_HIP(err);
    return RR_OK;
}

static rr_status ntt_forward_big(const NttTables &T, const NttFwdPack &pk, uint32_t ndesc, uint32_t logn,
                                 hipStream_t st) {
    const uint32_t l2 = big_l2(logn), l1 = logn - l2, N1 = 1u << l1, N2 = 1u << l2, TC = big_tc(logn);
    uint32_t qmax = 0;
    NttFwdPack rows;
    for (uint32_t i = 0; i < ndesc; i++) {
        qmax = std::max(qmax, pk.d[i].count);
        RR_CHECK((uint64_t)pk.d[i].count * N2 < (1ull << 32), RR_EINVAL, "too many transforms");
        rows.d[i] = NttFwdDesc{pk.d[i].dst, N1, 0, N1, pk.d[i].dst, pk.d[i].count * N2};
    }
    const ModP mp = ModP::make(T.prime);
    const size_t lds = (size_t)(N2 + 1) * TC * 4;
    hipLaunchKernelGGL(k_ntt_cols_fwd, dim3(qmax * (N1 / TC), ndesc), dim3(NTT_THREADS), lds, st, pk, logn,
                       T.tf, T.tfq, T.logn, T.prime, mp.inv);
    RR_HIP(hipGetLastError());
    return ntt_forward_pack(T, rows, ndesc, l1, st);
}

static rr_status ntt_inverse_big(const NttTables &T, const NttInvPack &pk, uint32_t ndesc, uint32_t logn,
                                 hipStream_t st) {
    const uint32_t l2 = big_l2(logn), l1 = logn - l2, N1 = 1u << l1, N2 = 1u << l2, TC = big_tc(logn);
    uint32_t qmax = 0;
    NttInvPack rows;
    for (uint32_t i = 0; i < ndesc; i++) {
        qmax = std::max(qmax, pk.d[i].count);
        RR_CHECK((uint64_t)pk.d[i].count * N2 < (1ull << 32), RR_EINVAL, "too many transforms");
        // rows in place: the spectra are the round's own scratch
        rows.d[i] = NttInvDesc{pk.d[i].src, pk.d[i].count * N2, 0, N1, const_cast<uint32_t *>(pk.d[i].src), N1};
    }
    rr_status s;
    if ((s = ntt_inverse_pack(T, rows, ndesc, l1, st)) != RR_OK) return s;
    const ModP mp = ModP::make(T.prime);
    const size_t lds = (size_t)(N2 + 1) * TC * 4;
    hipLaunchKernelGGL(k_ntt_cols_inv, dim3(qmax * (N1 / TC), ndesc), dim3(NTT_THREADS), lds, st, pk, logn,
                       T.ti, T.tiq, T.logn, T.prime, mp.inv);
    RR_HIP(hipGetLastError());
    return RR_OK;
}

rr_status ntt_forward_pack(const NttTables &T, const NttFwdPack &pk, uint32_t ndesc, uint32_t logn,
                           hipStream_t st) {
    RR_CHECK(logn >= 1 && logn <= T.logn && logn <= NTT_MAX_LOGN, RR_EINVAL, "NTT size 2^%u unsupported", logn);
    RR_CHECK(ndesc >= 1 && ndesc <= NTT_PACK, RR_EINVAL, "bad pack size");
    if (logn > NTT_LDS_LOGN) return ntt_forward_big(T, pk, ndesc, logn, st);
    rr_status e;
    if ((e = ntt_lds_attrs()) != RR_OK) return e;
    uint32_t bx = 0;
    for (uint32_t i = 0; i < ndesc; i++) bx = std::max(bx, blocks_for(pk.d[i].count, logn));
    hipLaunchKernelGGL(k_ntt_fwd, dim3(std::max(bx, 1u), ndesc), dim3(NTT_THREADS), lds_for(logn), st, pk, logn,
                       T.tf, T.tfq, T.logn, T.prime);
    RR_HIP(hipGetLastError());
    return RR_OK;
}

rr_status ntt_inverse_pack(const NttTables &T, const NttInvPack &pk, uint32_t ndesc, uint32_t logn,
                           hipStream_t st) {
    RR_CHECK(logn >= 1 && logn <= T.logn && logn <= NTT_MAX_LOGN, RR_EINVAL, "NTT size 2^%u unsupported", logn);
    RR_CHECK(ndesc >= 1 && ndesc <= NTT_PACK, RR_EINVAL, "bad pack size");
    if (logn > NTT_LDS_LOGN) return ntt_inverse_big(T, pk, ndesc, logn, st);
    rr_status e;
    if ((e = ntt_lds_attrs()) != RR_OK) return e;
    uint32_t bx = 0;
    for (uint32_t i = 0; i < ndesc; i++) bx = std::max(bx, blocks_for(pk.d[i].count, logn));
    hipLaunchKernelGGL(k_ntt_inv, dim3(std::max(bx, 1u), ndesc), dim3(NTT_THREADS), lds_for(logn), st, pk, logn,
                       T.ti, T.tiq, T.logn, T.prime);
    RR_HIP(hipGetLastError());
    return RR_OK;
}

rr_status pointwise_pack(const PwPack &pk, uint32_t ndesc, uint32_t N, hipStream_t st, uint32_t p) {
    RR_CHECK(ndesc >= 1 && ndesc <= NTT_PACK, RR_EINVAL, "bad pack size");
    RR_CHECK(N >= 1 && (N & (N - 1)) == 0, RR_EINVAL, "transform length %u is not a power of two", N);
    ModP mp = ModP::make(p);
    constexpr int TI = 4, TJ = 4;
    uint32_t rows = 0, cols = 0, logn = 0;
    for (uint32_t i = 0; i < ndesc; i++) {
        rows = std::max(rows, pk.d[i].rows);
        cols = std::max(cols, pk.d[i].cols);
    }
    while ((1u << logn) < N) logn++;
    const uint32_t rt = (rows + TI - 1) / TI, ct = (cols + TJ - 1) / TJ;
    const uint64_t threads = ((uint64_t)ndesc * rt * ct) << logn;
    RR_CHECK((threads + 255) / 256 < (1ull << 31), RR_EINVAL, "pointwise product too large");
    hipLaunchKernelGGL((k_pointwise<TI, TJ>), dim3((uint32_t)((threads + 255) / 256)), dim3(256), 0, st, pk, ndesc,
                       logn, rt, ct, p, mp.pd, mp.inv);
    RR_HIP(hipGetLastError());
    return RR_OK;
}

rr_status ntt_forward(const NttTables &T, const uint32_t *src, uint64_t sstride, uint32_t soff, uint32_t len,
                      uint32_t *dst, uint32_t count, uint32_t logn, hipStream_t st) {
    NttFwdPack pk;
    pk.d[0] = NttFwdDesc{src, sstride, soff, len, dst, count};
    return ntt_forward_pack(T, pk, 1, logn, st);
}

rr_status ntt_inverse(const NttTables &T, const uint32_t *src, uint32_t count, uint32_t logn, uint32_t lo,
                      uint32_t outlen, uint32_t *dst, uint64_t dstride, hipStream_t st) {
    NttInvPack pk;
    pk.d[0] = NttInvDesc{src, count, lo, outlen, dst, dstride};
    return ntt_inverse_pack(T, pk, 1, logn, st);
}

rr_status pointwise(const uint32_t *A, const uint32_t *B, uint32_t *C, uint32_t rows, uint32_t inner,
                    uint32_t cols, uint32_t N, uint32_t scale, hipStream_t st) {
    PwPack pk;
    pk.d[0] = PwDesc{A, B, C, rows, inner, cols, scale};
    return pointwise_pack(pk, 1, N, st);
}

}  // namespace rr

// ---------------------------------------------------------------------------
// rr_ntt_noalloc: rr_ntt's transforms on a caller-owned device buffer, with
// every work array in one fixed buffer per device, so a call allocates
// nothing (for callers under HIP graph capture or in loops that must not
// call hipMalloc).

namespace rr {

constexpr uint32_t NA_MAX_LOGN = 20;                      // transforms of up to 2^20 points
constexpr uint32_t NA_TW_WORDS = 4u << (NA_MAX_LOGN - 1);  // the four twiddle tables of 2^20-point transforms
constexpr uint32_t NA_MAX_WORDS = 1u << 22;                // largest count x 2^logn (16 MiB)
// This device's work buffer: twiddle tables (8 MiB), then the reordered
// input or the spectrum (16 MiB).  Host code uses it only through the
// address hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) uint32_t scratch[NA_TW_WORDS + NA_MAX_WORDS];

namespace {
struct NoAllocDev {
    uint32_t *ptr = nullptr;  // this device's copy of scratch
    bool ready = false;       // twiddle tables written
};
constexpr int NA_MAX_DEV = 64;
NoAllocDev g_na[NA_MAX_DEV];
std::mutex g_na_mu;
}  // namespace

// tf[k] = w^k, tfq[k] = floor(tf[k] 2^32 / p), and the same for w^-1, k < half
// (the tables NttTables::init builds on the host).
__global__ void k_ntt_twiddles(uint32_t *tf, uint32_t *tfq, uint32_t *ti, uint32_t *tiq, uint32_t half, uint32_t w,
                               uint32_t wi, uint32_t p, double inv) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= half) return;
    uint32_t a = 1, b = 1, x = w, y = wi;
    for (uint32_t e = k; e; e >>= 1) {
        if (e & 1) {
            a = mulmod(a, x, p, inv);
            b = mulmod(b, y, p, inv);
        }
        x = mulmod(x, x, p, inv);
        y = mulmod(y, y, p, inv);
    }
    tf[k] = a;
    tfq[k] = (uint32_t)(((uint64_t)a << 32) / p);
    ti[k] = b;
    tiq[k] = (uint32_t)(((uint64_t)b << 32) / p);
}

// dst[q][rev(i)] = src[q][i] * scale mod p, i < 2^logn, q < total / 2^logn.
__global__ void k_bitrev_scale(const uint32_t *src, uint32_t *dst, uint32_t logn, uint32_t total, uint32_t scale,
                               uint32_t p, double inv) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const uint32_t mask = (1u << logn) - 1, i = idx & mask;
    const uint32_t r = __brev(i) >> (32 - logn);
    const uint32_t v = src[idx];
    dst[(idx & ~mask) | r] = scale == 1 ? v : mulmod(v, scale, p, inv);
}

}  // namespace rr

extern "C" rr_status rr_ntt_noalloc(uint32_t *data, uint32_t logn, uint32_t count, int inverse, void *stream) {
    using namespace rr;
    clear_error();
    RR_CHECK(data && logn >= 1 && logn <= NA_MAX_LOGN && count >= 1, RR_EINVAL, "bad NTT arguments");
    RR_CHECK(((uint64_t)count << logn) <= NA_MAX_WORDS, RR_EINVAL,
             "count x 2^logn = %llu words exceeds the %u-word work buffer",
             (unsigned long long)((uint64_t)count << logn), NA_MAX_WORDS);
    hipStream_t st = (hipStream_t)stream;
    int dev = 0;
    RR_HIP(hipGetDevice(&dev));
    RR_CHECK(dev >= 0 && dev < NA_MAX_DEV, RR_EINVAL, "device %d out of range", dev);
    const uint32_t p = NTT_PRIME, half = 1u << (NA_MAX_LOGN - 1);
    const ModP mp = ModP::make(p);
    NoAllocDev &D = g_na[dev];
    {
        std::lock_guard<std::mutex> g(g_na_mu);
        if (!D.ready) {  // set-up, once per device: the buffer's address and the twiddle tables
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            RR_HIP(hipStreamIsCapturing(st, &cs));
            RR_CHECK(cs == hipStreamCaptureStatusNone, RR_ESTATE,
                     "the first rr_ntt_noalloc call on a device sets up its tables and must not be captured");
            if (!D.ptr) {
                void *addr = nullptr;
                RR_HIP(hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch)));
                D.ptr = (uint32_t *)addr;
            }
            const uint32_t w = h_powmod(NTT_ROOT, (p - 1) >> NA_MAX_LOGN, p), wi = h_invmod(w, p);
            hipLaunchKernelGGL(k_ntt_twiddles, dim3((half + 255) / 256), dim3(256), 0, st, D.ptr, D.ptr + half,
                               D.ptr + 2 * half, D.ptr + 3 * half, half, w, wi, p, mp.inv);
            RR_HIP(hipGetLastError());
            RR_HIP(hipStreamSynchronize(st));
            D.ready = true;
        }
    }
    NttTables T;  // a view of the tables in scratch (its own buffers stay empty)
    T.logn = NA_MAX_LOGN;
    T.tf = D.ptr;
    T.tfq = D.ptr + half;
    T.ti = D.ptr + 2 * half;
    T.tiq = D.ptr + 3 * half;
    uint32_t *work = D.ptr + NA_TW_WORDS;
    const uint32_t N = 1u << logn, total = count << logn;
    const dim3 pg((total + 255) / 256), pt(256);
    rr_status s;
    if (!inverse) {  // spectrum (bit-reversed) into work, then into data in natural order
        if ((s = ntt_forward(T, data, N, 0, N, work, count, logn, st)) != RR_OK) return s;
        hipLaunchKernelGGL(k_bitrev_scale, pg, pt, 0, st, work, data, logn, total, 1u, p, mp.inv);
    } else {  // bit-reversed and scaled by 1/N into work, then the inverse transform into data
        hipLaunchKernelGGL(k_bitrev_scale, pg, pt, 0, st, data, work, logn, total, h_invmod(N, p), p, mp.inv);
        RR_HIP(hipGetLastError());
        if ((s = ntt_inverse(T, work, count, logn, 0, N, data, N, st)) != RR_OK) return s;
    }
    RR_HIP(hipGetLastError());
    return RR_OK;
}
