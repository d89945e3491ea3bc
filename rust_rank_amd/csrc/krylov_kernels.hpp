// krylov_kernels.hpp -- device building blocks of the fused sparse pass:
// one row of the mod-p sparse product on B lanes, and the per-lane
// accumulation of the symmetric b x b projection.
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
    uint32_t col[64];  // column indices of the current chunk, per row slot
    uint32_t val[64];  // values of the current chunk (general matrices)
    uint32_t o[64];    // produced row entries, for the projection
};

__device__ __forceinline__ void wave_sync() { __builtin_amdgcn_wave_barrier(); }

// One row of the sparse product on the B lanes of a row slot; lane j returns
// sum_e v_e * in[col_e][j] mod p.  The slot loads up to B column indices (and
// values) with one coalesced load, shares them through LDS and issues up to
// 16 independent gathers before using any, so a row costs about two dependent
// memory latencies per B nonzeros.  Lanes of one slot share the row, so they
// run the same trip counts.
//   packed:  |v| <= 8, x < 2^31, each term < 2^34 (exact in fp64), folded per chunk
//   general: v = vh 2^16 + vl, vh x < 2^46, vl x < 2^47, <= 64 terms per chunk
template <int B, bool PACKED>
__device__ __forceinline__ uint32_t row_dot(const uint32_t *__restrict__ col, const uint32_t *__restrict__ val,
                                            const uint32_t *__restrict__ in, uint32_t beg, uint32_t end, int j,
                                            WaveLds &w, double pd, double inv) {
    constexpr int GS = B < 16 ? B : 16;
    const int lane = (int)(threadIdx.x & 63), lb = lane & ~(B - 1);
    double a0 = 0.0, a1 = 0.0;
    for (uint32_t c0 = beg; c0 < end; c0 += B) {
        const uint32_t n = min((uint32_t)B, end - c0);
        w.col[lane] = (uint32_t)j < n ? col[c0 + j] : 0u;
        if constexpr (!PACKED) w.val[lane] = (uint32_t)j < n ? val[c0 + j] : 0u;
        wave_sync();
        for (uint32_t e = 0; e < n; e += GS) {
            uint32_t c[GS], x[GS], v[GS];
#pragma unroll
            for (int k = 0; k < GS; k++) {
                c[k] = w.col[lb + (int)((e + k) & (B - 1))];
                if constexpr (!PACKED) v[k] = w.val[lb + (int)((e + k) & (B - 1))];
            }
#pragma unroll
            for (int k = 0; k < GS; k++) {
                const uint32_t idx = PACKED ? (c[k] & 0x0FFFFFFFu) : c[k];
                x[k] = (e + k < n) ? in[(size_t)idx * B + j] : 0u;
            }
#pragma unroll
            for (int k = 0; k < GS; k++) {
                if constexpr (PACKED) {
                    const double t = (double)((int32_t)c[k] >> 28);
                    if (k & 1) a1 = fma(t, (double)x[k], a1);
                    else a0 = fma(t, (double)x[k], a0);
                } else {
                    const double xd = (double)x[k];
                    a0 = fma((double)(v[k] >> 16), xd, a0);
                    a1 = fma((double)(v[k] & 0xFFFFu), xd, a1);
                }
            }
        }
        wave_sync();
        a0 = fold(a0, pd, inv);
        a1 = fold(a1, pd, inv);
    }
    if constexpr (PACKED) return (uint32_t)fold(a0 + a1, pd, inv);
    else return (uint32_t)fold(fma(a0, 65536.0, a1), pd, inv);
}

// ---------------------------------------------------------------------------
// Per-lane accumulation of the symmetric projection sum_r o_r s_r^T.
// Lane j of a row pairs its own s_j with o_{(j+k) mod B} for k = 0..B/2 (read
// from LDS), which covers every unordered pair once (k = B/2 only for
// j < B/2).  s is split in 16-bit halves so each product fits fp64 exactly;
// partial sums fold every 64 rows.
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

}  // namespace rr
