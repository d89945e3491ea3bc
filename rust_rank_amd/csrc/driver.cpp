// driver.cpp -- the rank driver over the C ABI: Krylov sequence on the device,
// device PM-basis, rank formula.
//
// Replaces the reference's rank driver (src/main.rs:397-606): draws are the
// caller's (src/main.rs:464-473 uses an unseeded RNG, so they are inputs
// here), the sequence length and PM-basis order follow either the
// reference's policy (src/main.rs:497-504, 599-603: next power of two of
// floor(2 min(m,n)/b + 4), plus one entry; largest power of two <= length) or
// the exact policy used by the workloads (length 2 min(m,n)/b + 8, order
// length - 1), and the rank is returned instead of printed
// (src/blockbmtest.rs:598-599).
#include "common.hpp"
#include "batch.hpp"
#include <vector>
#include <exception>
#include <chrono>


extern "C" {

uint64_t rr_seq_len(uint64_t m, uint64_t n, uint32_t b, int reference_policy) {
    if (b == 0) return 0;
    uint64_t mn = m < n ? m : n;
    if (!reference_policy) return 2 * mn / b + 8;
    // src/main.rs:497-504: floor(2 d / num_v + 4), then next power of two (num_v > 1),
    // and the loop produces max_nlen + 1 entries (src/main.rs:109-131)
    uint64_t len = (uint64_t)(2.0 * (double)mn / (double)b + 4.0);
    if (b > 1) {
        uint64_t q = 1;
        while (q < len) q <<= 1;
        len = q;
    }
    return len + 1;
}

uint64_t rr_pm_order(uint64_t seq_len, int reference_policy) {
    if (seq_len < 2) return 0;
    if (!reference_policy) return seq_len - 1;
    uint64_t d = 1;  // src/main.rs:599-603: largest power of two <= len
    while (d * 2 <= seq_len) d *= 2;
    if (d >= seq_len) d /= 2;
    return d;
}

rr_status rr_rank(const rr_csr *A, uint32_t p, uint32_t b, const uint32_t *row_precond, const uint32_t *col_precond,
                  const uint32_t *V, uint64_t seq_len, uint64_t order, int64_t *rank_out, double *ms_out) {
    return rr::guarded([&]() -> rr_status {
        rr::clear_error();
        RR_CHECK(A && rank_out, RR_EINVAL, "null argument");
        auto t0 = std::chrono::steady_clock::now();
        if (seq_len == 0) seq_len = rr_seq_len(A->m, A->n, b, 0);
        if (order == 0) order = rr_pm_order(seq_len, 0);
        RR_CHECK(order >= 1 && order < seq_len, RR_EINVAL, "order must satisfy 1 <= order < seq_len");
        const uint64_t P = (uint64_t)b * (b + 1) / 2;
        std::vector<uint32_t> seq;
        try {
            seq.resize(seq_len * P);
        } catch (const std::exception &) {
            rr::set_error("host allocation of %llu sequence words failed", (unsigned long long)(seq_len * P));
            return RR_ENOMEM;
        }
        rr_krylov *k = nullptr;
        rr_status s = rr_krylov_create(A, p, b, row_precond, col_precond, V, 0, &k);
        if (s != RR_OK) return s;
        if (order >= rr::PM_STREAM_D && A->row_ptr[A->m] >= rr::RANK_STREAM_NNZ) {
            // the PM-basis runs on the leading coefficients while the sequence grows
            s = rr::rank_streamed(k, b, seq_len, order, p, seq.data(), rank_out);
        } else {
            s = rr_krylov_run(k, seq_len, seq.data());
            if (s == RR_OK)  // the sequence is still on the device: G, PM-basis and rank stay there
                s = rr::rank_from_device_seq(rr::krylov_device_seq(k), b, order, p, rr::krylov_stream(k), rank_out);
        }
        rr_krylov_destroy(k);
        if (s != RR_OK) return s;
        if (ms_out)
            *ms_out = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return RR_OK;
    });
}

// The reference's analysis of a stored sequence (src/main.rs:397-401,
// 423-436, 599-606: a file name ending in .wdm loads the checkpoint and runs
// PM_Basis and analyze_pm_basis on its sequence).  The sequence goes to the
// device once; G, the PM-basis and the rank inputs stay there.
rr_status rr_rank_from_wdm(const char *path, int reference_policy, uint64_t order, int64_t *rank_out) {
    return rr::guarded([&]() -> rr_status {
        rr::clear_error();
        RR_CHECK(path && rank_out, RR_EINVAL, "null argument");
        uint64_t m = 0, n = 0, len = 0;
        uint32_t p = 0, b = 0;
        rr_status s;
        if ((s = rr_wdm_info(path, &m, &n, &p, &len, &b)) != RR_OK) return s;
        RR_CHECK(b >= 1 && b <= 64, RR_EINVAL, "%s: b=%u must be in 1..64", path, b);
        RR_CHECK(p > 2 && p < (1u << 31) && rr::h_is_prime(p), RR_EINVAL, "%s: p=%u must be a prime in (2, 2^31)",
                 path, p);
        if (order == 0) order = rr_pm_order(len, reference_policy);
        RR_CHECK(order >= 1 && order < len, RR_EINVAL, "%s: order %llu needs 1 <= order < %llu entries", path,
                 (unsigned long long)order, (unsigned long long)len);
        const uint64_t P = (uint64_t)b * (b + 1) / 2;
        std::vector<uint32_t> rp, cp, V, curv, seq;
        try {
            rp.resize(m);
            cp.resize(n);
            V.resize(n * b);
            curv.resize(n * b);
            seq.resize(len * P);
        } catch (const std::exception &) {
            rr::set_error("%s: host allocation for a %llux%llu, b=%u, %llu-entry file failed", path,
                          (unsigned long long)m, (unsigned long long)n, b, (unsigned long long)len);
            return RR_ENOMEM;
        }
        if ((s = rr_wdm_load(path, rp.data(), cp.data(), V.data(), curv.data(), seq.data())) != RR_OK) return s;
        for (size_t i = 0; i < (order + 1) * P; i++)
            RR_CHECK(seq[i] < p, RR_EINVAL, "%s: sequence entry %zu >= p", path, i);
        hipStream_t st;
        RR_HIP(rr::stream_create(&st));
        struct SG {
            hipStream_t s;
            ~SG() { rr::stream_destroy(s); }
        } guard{st};
        rr::DevBuf dseq;
        if ((s = dseq.alloc((order + 1) * P * 4)) != RR_OK) return s;
        RR_HIP(hipMemcpyAsync(dseq.ptr, seq.data(), (order + 1) * P * 4, hipMemcpyHostToDevice, st));
        return rr::rank_from_device_seq(dseq.as<uint32_t>(), b, order, p, st, rank_out);
    });
}

}  // extern "C"
