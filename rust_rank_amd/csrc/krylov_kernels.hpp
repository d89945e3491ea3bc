// krylov_kernels.hpp -- device building blocks of the fused sparse pass:
// one row of the mod-p sparse product on B lanes, and the per-lane
// accumulation of the symmetric b x b projection.
#pragma once
#include "common.hpp"

namespace rr {

constexpr int WG = 256;  // threads per workgroup (4 wavefronts)

// ---------------------------------------------------------------------------
// one row of the sparse product, lane j: sum_e v_e * in[col_e][j] mod p
template <int B, bool PACKED>
__device__ __forceinline__ uint32_t row_dot(const uint32_t *__restrict__ col, const uint32_t *__restrict__ val,
                                            const uint32_t *__restrict__ in, uint32_t beg, uint32_t end, int j,
                                            double pd, double inv) {
    if constexpr (PACKED) {
        // |v| <= 8, x < 2^31: each term < 2^34, exact in fp64 for 2^19 terms
        double acc0 = 0.0, acc1 = 0.0;
        uint32_t e = beg;
        for (; e + 4 <= end; e += 4) {
            uint32_t c0 = col[e], c1 = col[e + 1], c2 = col[e + 2], c3 = col[e + 3];
            uint32_t x0 = in[(size_t)(c0 & 0x0FFFFFFFu) * B + j];
            uint32_t x1 = in[(size_t)(c1 & 0x0FFFFFFFu) * B + j];
            uint32_t x2 = in[(size_t)(c2 & 0x0FFFFFFFu) * B + j];
            uint32_t x3 = in[(size_t)(c3 & 0x0FFFFFFFu) * B + j];
            acc0 = fma((double)((int32_t)c0 >> 28), (double)x0, acc0);
            acc1 = fma((double)((int32_t)c1 >> 28), (double)x1, acc1);
            acc0 = fma((double)((int32_t)c2 >> 28), (double)x2, acc0);
            acc1 = fma((double)((int32_t)c3 >> 28), (double)x3, acc1);
        }
        for (; e < end; e++) {
            uint32_t c0 = col[e];
            uint32_t x0 = in[(size_t)(c0 & 0x0FFFFFFFu) * B + j];
            acc0 = fma((double)((int32_t)c0 >> 28), (double)x0, acc0);
        }
        return (uint32_t)fold(acc0 + acc1, pd, inv);
    } else {
        // v = vh 2^16 + vl: vh x < 2^46, vl x < 2^47; fold every 64 terms
        double hi = 0.0, lo = 0.0;
        uint32_t e = beg;
        while (e < end) {
            uint32_t stop = min(end, e + 64u);
            for (; e + 2 <= stop; e += 2) {
                uint32_t c0 = col[e], c1 = col[e + 1];
                uint32_t v0 = val[e], v1 = val[e + 1];
                double x0 = (double)in[(size_t)c0 * B + j];
                double x1 = (double)in[(size_t)c1 * B + j];
                hi = fma((double)(v0 >> 16), x0, hi);
                lo = fma((double)(v0 & 0xFFFFu), x0, lo);
                hi = fma((double)(v1 >> 16), x1, hi);
                lo = fma((double)(v1 & 0xFFFFu), x1, lo);
            }
            if (e < stop) {
                uint32_t c0 = col[e], v0 = val[e];
                double x0 = (double)in[(size_t)c0 * B + j];
                hi = fma((double)(v0 >> 16), x0, hi);
                lo = fma((double)(v0 & 0xFFFFu), x0, lo);
                e++;
            }
            hi = fold(hi, pd, inv);
            lo = fold(lo, pd, inv);
        }
        return (uint32_t)fold(fma(hi, 65536.0, lo), pd, inv);
    }
}

// ---------------------------------------------------------------------------
// Per-lane accumulation of the symmetric projection sum_r o_r s_r^T.
// Lane j of a row pairs its own s_j with o_{(j+k) mod B} for k = 0..B/2, which
// covers every unordered pair once (k = B/2 only for j < B/2).  s is split in
// 16-bit halves so each product fits fp64 exactly; partial sums fold every 64
// rows.
template <int B>
struct Gram {
    static constexpr int K = B / 2 + 1;
    double hi[K], lo[K];
    int rows;

    __device__ void init() {
#pragma unroll
        for (int k = 0; k < K; k++) hi[k] = lo[k] = 0.0;
        rows = 0;
    }
    // all 64 lanes must call this together
    __device__ __forceinline__ void add(uint32_t o, uint32_t s, int lane, double pd, double inv) {
        const int base = lane & ~(B - 1);
        const int j = lane & (B - 1);
        const double sh = (double)(s >> 16), sl = (double)(s & 0xFFFFu);
#pragma unroll
        for (int k = 0; k < K; k++) {
            uint32_t ok = (k == 0) ? o : (uint32_t)__shfl((int)o, base + ((j + k) & (B - 1)), 64);
            double od = (double)ok;
            hi[k] = fma(od, sh, hi[k]);
            lo[k] = fma(od, sl, lo[k]);
        }
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

}  // namespace rr
