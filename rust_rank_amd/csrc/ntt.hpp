// ntt.hpp -- device NTT and polynomial-matrix product primitives (mod 2013265921).
#pragma once
#include "common.hpp"
#include <vector>

namespace rr {

constexpr uint32_t NTT_PRIME = 2013265921u;  // 15 * 2^27 + 1 (src/ntt.rs:52)
constexpr uint32_t NTT_ROOT = 31u;           // generator (src/ntt.rs:53)
constexpr uint32_t NTT_LDS_LOGN = 14;        // one workgroup, one polynomial in 64 KB of LDS
constexpr uint32_t NTT_MAX_LOGN = 26;        // larger transforms: four steps through global memory

// Twiddle tables for transforms up to 2^logn points: w^k and w^-k for k < 2^(logn-1)
// with their Shoup companions.
struct NttTables {
    uint32_t logn = 0;
    uint32_t prime = NTT_PRIME, root = NTT_ROOT;  // the transforms run mod prime
    DevBuf fwd, fwdq, inv, invq;  // owned tables (init)
    // the tables the transforms read: the owned ones, or tables laid out by the caller
    const uint32_t *tf = nullptr, *tfq = nullptr, *ti = nullptr, *tiq = nullptr;
    rr_status init(uint32_t logn_max, hipStream_t stream, uint32_t prime = NTT_PRIME, uint32_t root = NTT_ROOT);
};

// Products mod a prime p that is not an NTT prime: three NTT primes with
// 2-adicity >= 26 (so every transform size up to 2^NTT_MAX_LOGN exists mod
// each) whose product, about 2^90.47, must exceed every exact coefficient of
// the product (checked per product: inner x terms x (p - 1)^2; for instance
// 128 x 2^21 terms of residues < 2^31, or any product of residues of a prime
// below 2^16 up to 2^26 points); each coefficient is recovered by Garner's
// method and reduced mod p.  CRT_PRIMES[0] is NTT_PRIME.
constexpr uint32_t CRT_PRIMES[3] = {2013265921u, 469762049u, 1811939329u};  // 15 2^27 + 1, 7 2^26 + 1, 27 2^26 + 1
constexpr uint32_t CRT_ROOTS[3] = {31u, 3u, 13u};                         // their generators
// dst[i] = src[(i / len) * sstride + soff + i % len] mod q, i < count * len
rr_status mod_copy(const uint32_t *src, uint64_t sstride, uint32_t soff, uint32_t len, uint32_t count, uint32_t q,
                   uint32_t *dst, hipStream_t st);
// C[i] = x mod p for the x < q0 q1 q2 with x = r_k[i] mod q_k (Garner), i < n
rr_status garner3(const uint32_t *r0, const uint32_t *r1, const uint32_t *r2, uint64_t n, uint32_t p, uint32_t *C,
                  hipStream_t st);

// Descriptor packs: one launch runs up to NTT_PACK independent transforms or
// pointwise products of the same size (blockIdx selects the descriptor).
constexpr uint32_t NTT_PACK = 96;
struct NttFwdDesc {  // dst[q][0..N) = spectrum of src[q * sstride + soff + k], k < len, q < count
    const uint32_t *src;
    uint64_t sstride;
    uint32_t soff, len;
    uint32_t *dst;
    uint32_t count;
};
struct NttInvDesc {  // dst[q * dstride + k] = coefficient lo + k of the inverse of src[q], k < outlen
    const uint32_t *src;
    uint32_t count, lo, outlen;
    uint32_t *dst;
    uint64_t dstride;
};
struct PwDesc {  // C[i][j][l] = scale * sum_r A[i][r][l] B[r][j][l]
    const uint32_t *A, *B;
    uint32_t *C;
    uint32_t rows, inner, cols, scale;
};
struct NttFwdPack {
    NttFwdDesc d[NTT_PACK];
};
struct NttInvPack {
    NttInvDesc d[NTT_PACK];
};
struct PwPack {
    PwDesc d[NTT_PACK];
};
rr_status ntt_forward_pack(const NttTables &T, const NttFwdPack &pk, uint32_t ndesc, uint32_t logn, hipStream_t st);
rr_status ntt_inverse_pack(const NttTables &T, const NttInvPack &pk, uint32_t ndesc, uint32_t logn, hipStream_t st);
rr_status pointwise_pack(const PwPack &pk, uint32_t ndesc, uint32_t N, hipStream_t st, uint32_t p = NTT_PRIME);

// dst[q][0..2^logn) = forward spectrum (bit-reversed order) of src[q * sstride + soff + k], k < len.
rr_status ntt_forward(const NttTables &T, const uint32_t *src, uint64_t sstride, uint32_t soff, uint32_t len,
                      uint32_t *dst, uint32_t count, uint32_t logn, hipStream_t st);
// dst[q * dstride + k] = coefficient lo + k (k < outlen) of the inverse of spectrum src[q] (no 1/N).
rr_status ntt_inverse(const NttTables &T, const uint32_t *src, uint32_t count, uint32_t logn, uint32_t lo,
                      uint32_t outlen, uint32_t *dst, uint64_t dstride, hipStream_t st);
// C[i][j][l] = scale * sum_r A[i][r][l] B[r][j][l], spectra of length N (mod 2013265921).
rr_status pointwise(const uint32_t *A, const uint32_t *B, uint32_t *C, uint32_t rows, uint32_t inner,
                    uint32_t cols, uint32_t N, uint32_t scale, hipStream_t st);
}  // namespace rr
