// krylov_kernels.hpp -- device building blocks of the fused sparse pass and
// the t = 0 projections: exact reductions mod p and the accumulation of the
// symmetric b x b projection.
#pragma once
#include "common.hpp"

namespace rr {

constexpr int WG = 256;  // threads per workgroup (4 wavefronts)

// ---------------------------------------------------------------------------
// Lanes of one wavefront exchange values through a per-wavefront LDS area
// (ds_bpermute shuffles measure ~10 lane-ops/clk/CU on MI355X, LDS reads 3x
// more).  LDS operations of one wavefront complete in order; the wave
// barrier keeps the compiler from moving them across each other.
struct WaveLds {
    uint32_t o[64];  // the rows' entries, for the projection
};

__device__ __forceinline__ void wave_sync() { __builtin_amdgcn_wave_barrier(); }

// ---------------------------------------------------------------------------
// Per-lane accumulation of the symmetric projection sum_r o_r s_r^T.
// Lane j of a row pairs its own s_j with o_{(j+k) mod B} for k = 0..B/2 (read
// from LDS), which covers every unordered pair once (k = B/2 only for
// j < B/2).  s is split in 16-bit halves so each product fits fp64 exactly;
// partial sums fold every 64 rows.
template <int B>
struct Gram {
    static constexpr int K = B / 2 + 1;
    static constexpr int LDS_WORDS = (WG / 64) * B * K;
    double hi[K], lo[K];
    int rows;

    __device__ void init() {
#pragma unroll
        for (int k = 0; k < K; k++) hi[k] = lo[k] = 0.0;
        rows = 0;
    }
    // all 64 lanes of the wavefront call this together
    __device__ __forceinline__ void add(uint32_t o, uint32_t s, int lane, WaveLds &w, double pd, double inv) {
        const int base = lane & ~(B - 1);
        const int j = lane & (B - 1);
        const double sh = (double)(s >> 16), sl = (double)(s & 0xFFFFu);
        w.o[lane] = o;
        wave_sync();
#pragma unroll
        for (int k = 0; k < K; k++) {
            const uint32_t ok = (k == 0) ? o : w.o[base + ((j + k) & (B - 1))];
            const double od = (double)ok;
            hi[k] = fma(od, sh, hi[k]);
            lo[k] = fma(od, sl, lo[k]);
        }
        wave_sync();
        if (++rows == 64) {
#pragma unroll
            for (int k = 0; k < K; k++) {
                hi[k] = fold(hi[k], pd, inv);
                lo[k] = fold(lo[k], pd, inv);
            }
            rows = 0;
        }
    }
    // Reduce over the wavefront's rows and the workgroup's wavefronts, then
    // write the b(b+1)/2 upper-triangle partial to out.
    __device__ void finish(uint32_t *out, uint32_t b, uint32_t p, double pd, double inv, uint32_t *lds) {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, j = lane & (B - 1);
        uint32_t v[K];
#pragma unroll
        for (int k = 0; k < K; k++) v[k] = (uint32_t)fold(fma(fold(hi[k], pd, inv), 65536.0, fold(lo[k], pd, inv)), pd, inv);
        // rows of the same wavefront: lanes j, j + B, j + 2B, ...
#pragma unroll
        for (int off = B; off < 64; off <<= 1) {
#pragma unroll
            for (int k = 0; k < K; k++) v[k] = addmod(v[k], (uint32_t)__shfl_xor((int)v[k], off, 64), p);
        }
        if (lane < B) {
#pragma unroll
            for (int k = 0; k < K; k++) lds[(wave * B + j) * K + k] = v[k];
        }
        __syncthreads();
        const int nw = blockDim.x >> 6;
        for (int t = threadIdx.x; t < B * K; t += blockDim.x) {
            int jj = t / K, k = t % K;
            uint32_t s = 0;
            for (int w = 0; w < nw; w++) s = addmod(s, lds[(w * B + jj) * K + k], p);
            int i1 = jj, i2 = (jj + k) & (B - 1);
            if (2 * k == B && jj >= B / 2) continue;  // that pair is covered by jj - B/2
            if ((uint32_t)i1 >= b || (uint32_t)i2 >= b) continue;
            uint32_t lo_i = min(i1, i2), hi_i = max(i1, i2);
            uint32_t ii = lo_i * b - lo_i * (lo_i - 1) / 2 + (hi_i - lo_i);
            out[ii] = s;
        }
    }
};

// x mod p for x < 2^63 (x = h 2^32 + l): fold h, multiply by 2^32 mod p, add l.
__device__ __forceinline__ uint32_t fold_u64(uint64_t x, uint32_t p, double pd, double inv) {
    const uint32_t h = (uint32_t)fold((double)(uint32_t)(x >> 32), pd, inv);
    const uint32_t c32 = (uint32_t)fold(4294967296.0, pd, inv);
    return (uint32_t)fold((double)mulmod(h, c32, p, inv) + (double)(uint32_t)x, pd, inv);
}

// ---------------------------------------------------------------------------
// The symmetric projection with the wavefront's 64 lanes sharing ONE copy of
// the b x b upper triangle.  Lane l owns column i = l % B and the ring offsets
// k in [h KS, (h + 1) KS), h = l / B, of the pairs (i, (i + k) mod B),
// k = 0 .. B/2 (k = B/2 only for i < B/2): every unordered pair exactly once.
// Rows are added one at a time from LDS: the producing lanes store each row's
// o twice in a row (2B words, so the ring needs no wrap-around) and s once.
// S[i][j] = sum_r s_r[i] o_r[j]; s is split in 16-bit halves, so each product
// is below 2^47 and 2^16 rows fit an unsigned 64-bit accumulator.
template <int B>
struct GramRing {
    static constexpr int H = 64 / B;            // lanes per column
    static constexpr int K = B / 2 + 1;         // ring offsets
    static constexpr int KS = (K + H - 1) / H;  // ring offsets per lane
    static constexpr int P = B * (B + 1) / 2;
    static constexpr int LDS_WORDS = (WG / 64) * P;
    uint64_t hi[KS], lo[KS];
    uint32_t rows;

    __device__ void init() {
#pragma unroll
        for (int t = 0; t < KS; t++) hi[t] = lo[t] = 0;
        rows = 0;
    }
    __device__ static __forceinline__ bool owns(int i, int k) { return k < K && !(2 * k == B && i >= B / 2); }
    // All 64 lanes call this together; o2: the row's o values twice, s_i =
    // s[i].  Accumulators of offsets a lane does not own collect products of
    // other pairs and are never written out, so the reads need no masking
    // (for B >= 8 they stay inside the doubled row).
    __device__ __forceinline__ void add(const uint32_t *o2, uint32_t s_i, uint32_t p, double pd, double inv) {
        const int lane = threadIdx.x & 63, i = lane % B, h = lane / B;
        const uint32_t sh = s_i >> 16, sl = s_i & 0xFFFFu;
        const uint32_t *src = o2 + i + h * KS;
#pragma unroll
        for (int t = 0; t < KS; t++) {
            uint32_t o;
            if constexpr (H * KS > B) o = o2[i + min(h * KS + t, K - 1)];
            else o = src[t];
            hi[t] += (uint64_t)o * sh;
            lo[t] += (uint64_t)o * sl;
        }
        if (++rows == 65536u) {
#pragma unroll
            for (int t = 0; t < KS; t++) {
                hi[t] = fold_u64(hi[t], p, pd, inv);
                lo[t] = fold_u64(lo[t], p, pd, inv);
            }
            rows = 0;
        }
    }
    // Sums the workgroup's wavefronts and writes the b(b+1)/2 upper-triangle
    // partial to out (row-major pair index).
    __device__ void finish(uint32_t *out, uint32_t b, uint32_t p, double pd, double inv, uint32_t *lds) {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, i = lane % B, h = lane / B;
        const uint32_t Pb = b * (b + 1) / 2;
#pragma unroll
        for (int t = 0; t < KS; t++) {
            const int k = h * KS + t;
            const int j = (i + k) % B;
            if (owns(i, k) && (uint32_t)i < b && (uint32_t)j < b) {
                const uint32_t v = addmod(mulmod(fold_u64(hi[t], p, pd, inv), 65536u, p, inv),
                                          fold_u64(lo[t], p, pd, inv), p);
                const uint32_t lo_i = (uint32_t)min(i, j), hi_i = (uint32_t)max(i, j);
                lds[wave * Pb + lo_i * b - lo_i * (lo_i - 1) / 2 + (hi_i - lo_i)] = v;
            }
        }
        __syncthreads();
        const int nw = blockDim.x >> 6;
        for (uint32_t t = threadIdx.x; t < Pb; t += blockDim.x) {
            uint32_t sum = 0;
            for (int w = 0; w < nw; w++) sum = addmod(sum, lds[w * Pb + t], p);
            out[t] = sum;
        }
    }
};

// ---------------------------------------------------------------------------
// The projection for lane width 32 on the matrix cores: S = sum_r o_r s_r^T
// as fp64 products v_mfma_f64_16x16x4f64 (lane l supplies A[l % 16][l / 16]
// and B[l / 16][l % 16]; register v of C holds C[4 v + l / 16][l % 16]) over
// four rows at a time, for the three upper 16 x 16 tiles (0,0), (0,1), (1,1)
// of the 32 x 32 triangle.  s is split into 16-bit halves, so every product
// is an integer below 2^47 and 64 rows of them stay below 2^53: the sums are
// exact whatever the order, and are folded mod p every 64 rows.
typedef double gram_v4d __attribute__((ext_vector_type(4)));

struct GramMma32 {
    static constexpr int P = 32 * 33 / 2;
    static constexpr int LDS_WORDS = (WG / 64) * P;
    gram_v4d acc[6];  // tile (0,0) hi, lo; (0,1) hi, lo; (1,1) hi, lo
    uint32_t rows;

    __device__ void init() {
#pragma unroll
        for (int t = 0; t < 6; t++) acc[t] = gram_v4d{0.0, 0.0, 0.0, 0.0};
        rows = 0;
    }
    // Rows k < nrows (<= 4) of o (stride ostride words) and s (stride 32 words).
    __device__ __forceinline__ void step(const uint32_t *o, uint32_t ostride, const uint32_t *s, uint32_t nrows,
                                         double pd, double inv) {
        const uint32_t lane = threadIdx.x & 63, k = lane >> 4, i = lane & 15;
        const bool ok = k < nrows;
        const uint32_t kk = ok ? k : 0;
        const uint32_t o0 = o[kk * ostride + i], o1 = o[kk * ostride + 16 + i];
        const uint32_t s0 = s[kk * 32 + i], s1 = s[kk * 32 + 16 + i];
        const double a0 = ok ? (double)o0 : 0.0, a1 = ok ? (double)o1 : 0.0;
        const double b0h = (double)(s0 >> 16), b0l = (double)(s0 & 0xFFFFu);
        const double b1h = (double)(s1 >> 16), b1l = (double)(s1 & 0xFFFFu);
        acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0h, acc[0], 0, 0, 0);
        acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0l, acc[1], 0, 0, 0);
        acc[2] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1h, acc[2], 0, 0, 0);
        acc[3] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1l, acc[3], 0, 0, 0);
        acc[4] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1h, acc[4], 0, 0, 0);
        acc[5] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1l, acc[5], 0, 0, 0);
        rows += 4;
        if (rows == 64) {
#pragma unroll
            for (int t = 0; t < 6; t++)
#pragma unroll
                for (int v = 0; v < 4; v++) acc[t][v] = fold(acc[t][v], pd, inv);
            rows = 0;
        }
    }
    // The eight rows of a slice: o rows doubled (stride 64 words), s rows (stride 32).
    __device__ __forceinline__ void add_slice(const uint32_t *o2, const uint32_t *s, double pd, double inv) {
        step(o2, 64, s, 4, pd, inv);
        step(o2 + 4 * 64, 64, s + 4 * 32, 4, pd, inv);
    }
    __device__ __forceinline__ void add_row(const uint32_t *o, const uint32_t *s, double pd, double inv) {
        step(o, 64, s, 1, pd, inv);
    }
    // Sums the workgroup's wavefronts and writes the b(b+1)/2 upper-triangle
    // partial to out (row-major pair index), as GramRing::finish.
    __device__ void finish(uint32_t *out, uint32_t b, uint32_t p, double pd, double inv, uint32_t *lds) {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const uint32_t Pb = b * (b + 1) / 2;
#pragma unroll
        for (int t = 0; t < 3; t++) {
            const uint32_t ti = t == 2 ? 1 : 0, tj = t == 0 ? 0 : 1;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const uint32_t i = 16 * ti + 4 * v + (lane >> 4), j = 16 * tj + (lane & 15);
                if (i <= j && j < b) {
                    const double h = fold(acc[2 * t][v], pd, inv), l = fold(acc[2 * t + 1][v], pd, inv);
                    lds[wave * Pb + i * b - i * (i - 1) / 2 + (j - i)] = (uint32_t)fold(fma(h, 65536.0, l), pd, inv);
                }
            }
        }
        __syncthreads();
        const int nw = blockDim.x >> 6;
        for (uint32_t t = threadIdx.x; t < Pb; t += blockDim.x) {
            uint32_t sum = 0;
            for (int w = 0; w < nw; w++) sum = addmod(sum, lds[w * Pb + t], p);
            out[t] = sum;
        }
    }
};

}  // namespace rr
