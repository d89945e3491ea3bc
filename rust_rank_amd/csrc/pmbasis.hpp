// pmbasis.hpp -- device PM-basis: jobs, the round scheduler and rank extraction.
#pragma once
#include "common.hpp"
#include "ntt.hpp"
#include <vector>
#include <utility>

namespace rr {

// Steps per leaf in batches and in the server: the in-leaf basis update
// costs m^2 rank per coefficient, and P and M of a leaf must fit in LDS, so
// wider blocks take shorter leaves (m = 64: 4, the most that fits).  With
// many jobs sharing rounds, long leaves mean fewer rounds and fewer products
// (m = 32: 16 steps made 512 batched matrices 10% faster than 8).
__host__ __device__ constexpr uint32_t leaf_order(uint32_t m) { return m <= 32 ? 16 : m <= 64 ? 4 : 1; }
// A PM-basis running alone is a chain of leaf launches; in batches and in
// the server the leaves of many jobs share rounds, and longer leaves mean
// fewer rounds.  Alone, the fastest orders measured were 16, 8, 4 for
// m = 8, 16, 32 and single steps from m = 64 on (order 11257 at m = 64:
// 733 ms against 784 with two steps and 893 with four).  The basis is the
// same either way.
__host__ __device__ constexpr uint32_t leaf_order_alone(uint32_t m) { return m <= 8 ? 16 : m <= 16 ? 8 : m <= 32 ? 4 : 1; }
// A single PM-basis fed beside the Krylov passes (its leftmost chain runs
// while the sequence grows, the rest after it): measured on the 200000 x
// 180000 matrix through rr_rank, 8 steps at m = 16 and 32 (b = 16: 1.19 s
// against 1.30 s with 16), 2 at m = 64 (1.13 s against 1.16 s with 4 and
// 1.21 s with 1).
__host__ __device__ constexpr uint32_t leaf_order_streamed(uint32_t m) { return m <= 32 ? 8 : m <= 64 ? 2 : 1; }

struct LeafDesc {
    const uint32_t *G;  // residual G(i, j, k) at G[(i * n + j) * gs + k], k < d
    uint64_t gs;
    int64_t *delta;
    uint32_t *M;  // [m][m][d + 1]
    uint32_t m, n, d;
};
// C [rows][cols][hi-lo] = window [lo, hi) of A [rows][inner][la] * B [inner][cols][ob .. ob+lb) (stride sb)
struct MulDesc {
    const uint32_t *A, *B;
    uint32_t *C;
    uint64_t sb;
    uint32_t la, lb, ob, lo, hi, rows, inner, cols;
};
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
        uint32_t *ptr = buf.as<uint32_t>() + top;
        top += words;
        return ptr;
    }
};

// An NTT product of a round: transform size, operands, and where its A
// spectrum is kept (keep) or its B spectrum taken from (given); both may be null.
struct NttProduct {
    uint32_t logn;
    MulDesc D;
    uint32_t *keep;
    const uint32_t *given;
};

struct Frame {
    const uint32_t *G;
    uint64_t gs, d;
    uint32_t *M;
    int stage;
    uint32_t *M1, *GG, *M2;
    size_t mark;
    // M1's spectrum, kept from the product M1 G for the product M2 M1 (both
    // transform at the same size): SP [m][m][2^sp_logn], valid when sp_logn >= 0
    uint32_t *SP = nullptr;
    int sp_logn = -1;
};

struct PmJob {
    uint32_t b = 0, m = 0;
    uint32_t K = 0;  // leaf order (0: chosen when the job joins)
    uint64_t d = 0;
    const uint32_t *G = nullptr;  // [m][b][d+1]
    uint64_t avail = ~0ull;       // leading coefficients of G written so far (the rest still to come)
    uint32_t *M = nullptr;        // [m][m][d+1]
    int64_t *delta = nullptr;     // device [m]
    Pool pool;
    std::vector<Frame> stack;
    bool done = false;
    // the next device operation, taken from the recursion but not yet launched
    bool has_op = false;
    int op_kind = 0;  // 0 leaf, 1 product
    LeafDesc op_leaf{};
    MulDesc op_mul{};
    Frame *op_keep = nullptr;   // product keeps its A spectrum in this frame's SP
    Frame *op_given = nullptr;  // product takes its B spectrum from this frame's SP
    uint64_t op_need = 0;       // leading coefficients of G the next operation reads
    rr_status err = RR_OK;      // set (with done) when this job's recursion failed on its own
    uint64_t tag = 0;           // the caller's name for the job (the server's request id)
};

struct PmScheduler {
    uint32_t p = 0;
    hipStream_t st = nullptr;
    NttTables tables;      // mod p when p is the NTT prime, else mod the first CRT prime
    NttTables crt_tabs[2];  // mod the second and third CRT primes (p not the NTT prime)
    DevBuf scratch;
    size_t scratch_words = 0, lds = 0;
    uint64_t rounds = 0, launches = 0;
    uint32_t waiting = 0;  // after a round: jobs whose next operation waits for coefficients of G
    // The jobs whose operations the last round took.  A round that fails
    // before taking any (its work space could not be allocated) leaves every
    // job as it was and this list empty; one that fails after taking them
    // (a launch failed) leaves those jobs unusable.
    std::vector<PmJob *> last;
    // tests of failure handling: a round in which the job tagged so has an
    // operation ready fails with RR_ENOMEM before taking any (0: never)
    uint64_t fail_tag = 0;

    rr_status ensure_scratch(size_t words);
    rr_status init(uint32_t p, hipStream_t st, uint32_t max_logn, uint32_t max_m, uint32_t max_n);
    rr_status add(PmJob *J);                                          // job joins; its delta is zeroed
    rr_status round(std::vector<PmJob *> &jobs, bool *launched);     // next products, else next leaves
    rr_status run(std::vector<PmJob *> &jobs);                        // init + add + rounds to completion
    rr_status init_tables(uint32_t logn);  // the transform tables p's products need, up to 2^logn points
    rr_status run_ntt_products(std::vector<NttProduct> &ops);
    rr_status run_crt_products(std::vector<NttProduct> &ops);  // p not the NTT prime: three primes + Garner
};

rr_status pm_basis_batch(uint32_t p, std::vector<PmJob *> &jobs, hipStream_t st, uint64_t *rounds,
                         uint64_t *launches);
rr_status pm_basis_device(uint32_t p, uint32_t b, uint64_t d, const uint32_t *G, uint32_t *M, int64_t *delta_host,
                          hipStream_t st);
rr_status build_G_device(const uint32_t *seq_dev, uint32_t b, uint64_t len, uint32_t *G, hipStream_t st);
// coefficients [k0, k1) of the same G (len = its coefficients per entry);
// seq_dev holds the sequence from entry s0 <= k0 on
rr_status build_G_range(const uint32_t *seq_dev, uint32_t b, uint64_t len, uint64_t k0, uint64_t k1, uint32_t *G,
                        hipStream_t st, uint64_t s0 = 0);
int64_t rank_from_basis_host(const uint32_t *basis, const int64_t *delta, uint32_t b, uint64_t len, uint32_t p,
                             rr_status *st);

}  // namespace rr
