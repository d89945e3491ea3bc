// batch.hip -- many rank problems on one GPU: batched Krylov ticks over a
// tile list, batched PM-basis rounds, on-device rank extraction.
//
// The reference solves one matrix per process (src/main.rs:254-639); this is
// the MI355X-native throughput path for many small and medium matrices.
// Every request keeps its matrix, vectors and partial sequence resident in
// HBM.  A tick advances every active request by one sequence entry: each
// request contributes the tiles of its next pass (by A, by A^T, or the t = 0
// projection), tiles of all requests run in one launch per lane width, and a
// reduction kernel turns each request's tile partials into its S_t.  Requests
// never share a reduction, so each result is bit-identical alone and in any
// mix.  When a sequence is complete its PM-basis joins the batched scheduler
// (one launch per operation kind per round for all requests), and the rank
// is computed from the two coefficient matrices the formula needs, extracted
// on the device.
#include "common.hpp"
#include "krylov_kernels.hpp"
#include "pmbasis.hpp"
#include "batch.hpp"
#include <vector>
#include <algorithm>
#include <cstring>
#include <chrono>
#include <memory>
#include <cmath>

namespace rr {

// ---------------------------------------------------------------------------
// kernels

template <int B>
__global__ __launch_bounds__(WG) void k_tiles(const Tile *__restrict__ tiles, uint32_t ntiles, uint32_t p, double pd,
                                              double inv) {
    constexpr int G = 64 / B;
    __shared__ uint32_t lds[Gram<B>::LDS_WORDS];
    __shared__ WaveLds wl[WG / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int slot = lane / B, j = lane & (B - 1);
    WaveLds &w = wl[wave];
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile T = tiles[t];
        Gram<B> g;
        g.init();
        for (uint32_t base = T.row_beg + wave * G; base < T.row_end; base += (WG / 64) * G) {
            const uint32_t r = base + slot;
            uint32_t o = 0, s = 0;
            if (r < T.row_end) {
                if (T.mode == TILE_GRAM) {
                    o = s = T.in[(size_t)r * B + j];
                } else {
                    const uint32_t beg = T.rowptr[r], end = T.rowptr[r + 1];
                    o = T.mode == TILE_PACKED ? row_dot<B, true>(T.col, T.val, T.in, beg, end, j, w, pd, inv)
                                              : row_dot<B, false>(T.col, T.val, T.in, beg, end, j, w, pd, inv);
                    s = mulmod(o, T.scale[r], p, inv);
                    T.out[(size_t)r * B + j] = s;
                }
            }
            g.add(o, s, lane, w, pd, inv);
        }
        g.finish(T.part, T.b, p, pd, inv, lds);
        __syncthreads();  // lds is reused by the next tile
    }
}

// out[ii] = sum_k part[k * P + ii] mod p, one workgroup per job
__global__ void k_reduce_jobs(const RedJob *__restrict__ jobs, uint32_t p) {
    const RedJob J = jobs[blockIdx.x];
    for (uint32_t ii = threadIdx.x; ii < J.P; ii += blockDim.x) {
        uint64_t s = 0;
        for (uint32_t k = 0; k < J.ntiles; k++) s += J.part[(size_t)k * J.P + ii];
        J.out[ii] = (uint32_t)(s % p);
    }
}

// Inputs of the rank formula (src/sigma_basis.rs:115-145, src/blockbmtest.rs:598) from a
// basis [2b][2b][len]: out[0] = K (coefficients of the reversed generator),
// out[1 .. b^2] = F_0, out[1 + b^2 .. 2b^2] = F_{K-1}.
__global__ void k_rank_extract(const uint32_t *basis, const int64_t *delta, uint32_t b, uint64_t len, uint32_t *out) {
    __shared__ uint32_t maxk;
    __shared__ int64_t del[64];
    const uint32_t n2 = 2 * b;
    if (threadIdx.x == 0) maxk = 0;
    for (uint32_t i = threadIdx.x; i < b; i += blockDim.x) del[i] = -delta[i];
    __syncthreads();
    for (uint32_t ij = threadIdx.x; ij < b * b; ij += blockDim.x) {
        uint32_t i = ij / b, j = ij % b;
        int64_t dl = del[i];
        if (dl < 0 || (uint64_t)dl >= len) continue;
        const uint32_t *q = basis + ((size_t)i * n2 + j) * len;
        for (int64_t c = 0; c <= dl; c++)
            if (q[c] != 0) {
                atomicMax(&maxk, (uint32_t)(dl - c));
                break;
            }
    }
    __syncthreads();
    const uint32_t K = maxk + 1;
    if (threadIdx.x == 0) out[0] = K;
    for (uint32_t ij = threadIdx.x; ij < b * b; ij += blockDim.x) {
        uint32_t i = ij / b, j = ij % b;
        int64_t dl = del[i];
        const uint32_t *q = basis + ((size_t)i * n2 + j) * len;
        bool ok = dl >= 0 && (uint64_t)dl < len;
        out[1 + ij] = ok ? q[dl] : 0;
        int64_t c = dl - (int64_t)(K - 1);
        out[1 + b * b + ij] = (ok && c >= 0) ? q[c] : 0;
    }
}

__global__ void k_scale_rows_b(const uint32_t *in, const uint32_t *d, uint32_t *out, uint32_t nrows, uint32_t B,
                               uint32_t p, double inv) {
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)nrows * B) return;
    out[idx] = mulmod(in[idx], d[idx / B], p, inv);
}

// ---------------------------------------------------------------------------
// host side

template <typename T>
static rr_status upload_vec(DevBuf &buf, const std::vector<T> &v, hipStream_t st) {
    rr_status s = buf.alloc(v.size() * sizeof(T));
    if (s != RR_OK) return s;
    if (!v.empty()) RR_HIP(hipMemcpyAsync(buf.ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
    return RR_OK;
}

static uint32_t host_rank_bxb(std::vector<uint32_t> a, uint32_t b, uint32_t p) {
    uint32_t rank = 0;
    for (uint32_t c = 0; c < b && rank < b; c++) {
        uint32_t piv = rank;
        while (piv < b && a[(size_t)piv * b + c] == 0) piv++;
        if (piv == b) continue;
        for (uint32_t j = 0; j < b; j++) std::swap(a[(size_t)rank * b + j], a[(size_t)piv * b + j]);
        uint32_t iv = h_invmod(a[(size_t)rank * b + c], p);
        for (uint32_t j = c; j < b; j++) a[(size_t)rank * b + j] = h_mulmod(a[(size_t)rank * b + j], iv, p);
        for (uint32_t r = rank + 1; r < b; r++) {
            uint32_t f = a[(size_t)r * b + c];
            if (!f) continue;
            for (uint32_t j = c; j < b; j++) {
                uint32_t sub = h_mulmod(f, a[(size_t)rank * b + j], p);
                uint32_t &x = a[(size_t)r * b + j];
                x = x >= sub ? x - sub : x + p - sub;
            }
        }
        rank++;
    }
    return rank;
}

uint64_t request_bytes(uint64_t m, uint64_t n, uint64_t nnz, uint32_t b, uint64_t L, uint64_t d, bool packed) {
    uint64_t B = 1;
    while (B < b) B <<= 1;
    uint64_t P = (uint64_t)b * (b + 1) / 2;
    uint64_t words = 2 * nnz + (m + 1) + (n + 1) + (packed ? 0 : 2 * nnz) + m + 2 * n  // operands + scales
                     + B * (2 * n + m)                                                   // z, x0, y
                     + L * P                                                             // sequence
                     + (uint64_t)2 * b * b * (d + 1) + (uint64_t)4 * b * b * (d + 1)     // G, M
                     + (uint64_t)2 * (4ull * b * b * (d / 2 + 2) + 2ull * b * b * (d / 2 + 2)) + 2 * b * b + 64;
    return words * 4;
}

uint64_t Request::device_bytes() const { return request_bytes(m, n, nnz, b, L, d, packed); }

rr_status Request::init(const rr_csr *A, uint32_t p_, uint32_t b_, const uint32_t *dr, const uint32_t *dcv,
                        const uint32_t *V, uint64_t L_, uint64_t d_, uint32_t tile_rows, hipStream_t st) {
    p = p_;
    b = b_;
    B = 1;
    while (B < b) B <<= 1;
    P = b * (b + 1) / 2;
    m = A->m;
    n = A->n;
    L = L_;
    d = d_;
    RR_CHECK(p > 2 && p < (1u << 31) && h_is_prime(p), RR_EINVAL, "p=%u must be a prime in (2, 2^31)", p);
    RR_CHECK(b >= 1 && b <= 64, RR_EINVAL, "b=%u must be in 1..64", b);
    RR_CHECK(m >= 1 && n >= 1 && m < (1ull << 31) && n < (1ull << 31), RR_EINVAL, "matrix size out of range");
    RR_CHECK(L >= 2 && d >= 1 && d < L, RR_EINVAL, "need 1 <= d < L (L=%llu d=%llu)", (unsigned long long)L,
             (unsigned long long)d);
    RR_CHECK(p != NTT_PRIME || d + 1 <= (1ull << NTT_MAX_LOGN), RR_EINVAL, "order above the transform limit");
    RR_CHECK(A->row_ptr && A->row_ptr[0] == 0, RR_EINVAL, "row_ptr[0] must be 0");
    for (uint64_t r = 0; r < m; r++) RR_CHECK(A->row_ptr[r + 1] >= A->row_ptr[r], RR_EINVAL, "row_ptr not monotone");
    nnz = A->row_ptr[m];
    RR_CHECK(nnz < (1ull << 32), RR_EINVAL, "nnz must be < 2^32");
    packed = m < (1ull << 28) && n < (1ull << 28);
    for (uint64_t e = 0; e < nnz; e++) {
        RR_CHECK(A->col[e] < n && A->val[e] < p, RR_EINVAL, "entry %llu out of range", (unsigned long long)e);
        uint32_t v = A->val[e];
        if (!(v <= 7 || v >= p - 8)) packed = false;
    }
    for (uint64_t i = 0; i < m; i++) RR_CHECK(dr[i] >= 1 && dr[i] < p, RR_EINVAL, "row_precond out of range");
    for (uint64_t i = 0; i < n; i++) RR_CHECK(dcv[i] >= 1 && dcv[i] < p, RR_EINVAL, "col_precond out of range");
    for (uint64_t i = 0; i < n * b; i++) RR_CHECK(V[i] < p, RR_EINVAL, "V out of range");

    // A and A^T, raw values (src/matrices.rs:734-778 transpose order)
    std::vector<uint32_t> rp(m + 1), cw(nnz), vw(packed ? 0 : nnz), rpt(n + 1, 0), ct(nnz), vt(packed ? 0 : nnz);
    for (uint64_t r = 0; r <= m; r++) rp[r] = (uint32_t)A->row_ptr[r];
    for (uint64_t e = 0; e < nnz; e++) rpt[A->col[e] + 1]++;
    for (uint64_t i = 0; i < n; i++) rpt[i + 1] += rpt[i];
    std::vector<uint32_t> next(rpt.begin(), rpt.end() - 1);
    for (uint64_t r = 0; r < m; r++)
        for (uint64_t e = rp[r]; e < rp[r + 1]; e++) {
            uint32_t c = A->col[e], v = A->val[e], pos = next[c]++;
            if (packed) {
                int32_t sv = v <= 7 ? (int32_t)v : (int32_t)v - (int32_t)p;
                uint32_t code = ((uint32_t)sv & 0xFu) << 28;
                cw[e] = c | code;
                ct[pos] = (uint32_t)r | code;
            } else {
                cw[e] = c;
                vw[e] = v;
                ct[pos] = (uint32_t)r;
                vt[pos] = v;
            }
        }
    std::vector<uint32_t> dc2(n);
    for (uint64_t i = 0; i < n; i++) dc2[i] = h_mulmod(dcv[i], dcv[i], p);
    std::vector<uint32_t> xpad((size_t)n * B, 0);
    for (uint64_t r = 0; r < n; r++) std::memcpy(&xpad[r * B], &V[r * b], b * 4);
    rr_status s;
    if ((s = upload_vec(a_rp, rp, st)) != RR_OK) return s;
    if ((s = upload_vec(a_col, cw, st)) != RR_OK) return s;
    if ((s = upload_vec(a_val, vw, st)) != RR_OK) return s;
    if ((s = upload_vec(t_rp, rpt, st)) != RR_OK) return s;
    if ((s = upload_vec(t_col, ct, st)) != RR_OK) return s;
    if ((s = upload_vec(t_val, vt, st)) != RR_OK) return s;
    if ((s = upload_vec(dr_d, std::vector<uint32_t>(dr, dr + m), st)) != RR_OK) return s;
    if ((s = upload_vec(dc_d, std::vector<uint32_t>(dcv, dcv + n), st)) != RR_OK) return s;
    if ((s = upload_vec(dc2_d, dc2, st)) != RR_OK) return s;
    if ((s = upload_vec(x0, xpad, st)) != RR_OK) return s;
    if ((s = z.alloc((size_t)n * B * 4)) != RR_OK) return s;
    if ((s = y.alloc((size_t)m * B * 4)) != RR_OK) return s;
    if ((s = seq.alloc((size_t)L * P * 4)) != RR_OK) return s;
    RR_HIP(hipMemsetAsync(y.ptr, 0, (size_t)m * B * 4, st));
    size_t tot = (size_t)n * B;
    ModP mp = ModP::make(p);
    hipLaunchKernelGGL(k_scale_rows_b, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, x0.as<uint32_t>(),
                       dc_d.as<uint32_t>(), z.as<uint32_t>(), (uint32_t)n, B, p, mp.inv);
    RR_HIP(hipGetLastError());

    // tiles of each pass kind
    auto make = [&](std::vector<Tile> &out, uint64_t rows, const uint32_t *rowptr, const uint32_t *col,
                    const uint32_t *val, const uint32_t *in, uint32_t *o, const uint32_t *scale, uint32_t mode) {
        out.clear();
        for (uint64_t r0 = 0; r0 < rows; r0 += tile_rows) {
            Tile T;
            T.rowptr = rowptr;
            T.col = col;
            T.val = val;
            T.in = in;
            T.out = o;
            T.scale = scale;
            T.part = nullptr;
            T.row_beg = (uint32_t)r0;
            T.row_end = (uint32_t)std::min<uint64_t>(rows, r0 + tile_rows);
            T.b = b;
            T.mode = mode;
            out.push_back(T);
        }
    };
    uint32_t mode = packed ? TILE_PACKED : TILE_GENERAL;
    make(tiles_a, m, a_rp.as<uint32_t>(), a_col.as<uint32_t>(), a_val.as<uint32_t>(), z.as<uint32_t>(),
         y.as<uint32_t>(), dr_d.as<uint32_t>(), mode);
    make(tiles_at, n, t_rp.as<uint32_t>(), t_col.as<uint32_t>(), t_val.as<uint32_t>(), y.as<uint32_t>(),
         z.as<uint32_t>(), dc2_d.as<uint32_t>(), mode);
    make(tiles_gram, n, nullptr, nullptr, nullptr, x0.as<uint32_t>(), nullptr, nullptr, TILE_GRAM);
    t = 0;
    stage = STAGE_KRYLOV;
    return RR_OK;
}

const std::vector<Tile> &Request::next_tiles() const {
    if (t == 0) return tiles_gram;
    return (t & 1) ? tiles_a : tiles_at;
}

rr_status Request::start_pm(hipStream_t st) {
    rr_status s;
    if ((s = G.alloc((size_t)2 * b * b * (d + 1) * 4)) != RR_OK) return s;
    if ((s = M.alloc((size_t)4 * b * b * (d + 1) * 4)) != RR_OK) return s;
    if ((s = delta.alloc((size_t)2 * b * 8)) != RR_OK) return s;
    if ((s = rank_in.alloc((size_t)(2 * b * b + 1) * 4)) != RR_OK) return s;
    if ((s = build_G_device(seq.as<uint32_t>(), b, d + 1, G.as<uint32_t>(), st)) != RR_OK) return s;
    pm.b = b;
    pm.d = d;
    pm.G = G.as<uint32_t>();
    pm.M = M.as<uint32_t>();
    pm.delta = delta.as<int64_t>();
    stage = STAGE_PM;
    // vectors and operands are no longer needed
    a_rp.~DevBuf(); new (&a_rp) DevBuf();
    a_col.~DevBuf(); new (&a_col) DevBuf();
    a_val.~DevBuf(); new (&a_val) DevBuf();
    t_rp.~DevBuf(); new (&t_rp) DevBuf();
    t_col.~DevBuf(); new (&t_col) DevBuf();
    t_val.~DevBuf(); new (&t_val) DevBuf();
    return RR_OK;
}

// Rank from a device-resident basis [2b][2b][len] and delta: the two coefficient
// matrices the formula needs are extracted on the device (2b^2 + 1 words come back).
rr_status rank_from_device_basis(const uint32_t *M, const int64_t *delta, uint32_t b, uint64_t len, uint32_t p,
                                 uint32_t *scratch, hipStream_t st, int64_t *rank) {
    hipLaunchKernelGGL(k_rank_extract, dim3(1), dim3(256), 0, st, M, delta, b, len, scratch);
    RR_HIP(hipGetLastError());
    std::vector<uint32_t> h(2 * b * b + 1);
    RR_HIP(hipMemcpyAsync(h.data(), scratch, h.size() * 4, hipMemcpyDeviceToHost, st));
    RR_HIP(hipStreamSynchronize(st));
    uint64_t K = h[0];
    std::vector<uint32_t> F0(h.begin() + 1, h.begin() + 1 + b * b), FK(h.begin() + 1 + b * b, h.end());
    uint64_t r = (uint64_t)b * (K - 3) + host_rank_bxb(FK, b, p) + host_rank_bxb(F0, b, p);
    *rank = (int64_t)r;  // the reference's usize arithmetic wraps the same way when K < 3
    return RR_OK;
}

rr_status Request::finish_rank(hipStream_t st) {
    rr_status s = rank_from_device_basis(M.as<uint32_t>(), delta.as<int64_t>(), b, d + 1, p, rank_in.as<uint32_t>(),
                                         st, &rank);
    if (s != RR_OK) return s;
    stage = STAGE_DONE;
    return RR_OK;
}

// Sequence on the device [>= d+1][P] -> rank (G, PM-basis, extraction), all on the device.
rr_status rank_from_device_seq(const uint32_t *seq_dev, uint32_t b, uint64_t d, uint32_t p, hipStream_t st,
                               int64_t *rank) {
    DevBuf G, M, dl, sc;
    rr_status s;
    if ((s = G.alloc((size_t)2 * b * b * (d + 1) * 4)) != RR_OK) return s;
    if ((s = M.alloc((size_t)4 * b * b * (d + 1) * 4)) != RR_OK) return s;
    if ((s = dl.alloc((size_t)2 * b * 8)) != RR_OK) return s;
    if ((s = sc.alloc((size_t)(2 * b * b + 1) * 4)) != RR_OK) return s;
    if ((s = build_G_device(seq_dev, b, d + 1, G.as<uint32_t>(), st)) != RR_OK) return s;
    PmJob J;
    J.b = b;
    J.d = d;
    J.G = G.as<uint32_t>();
    J.M = M.as<uint32_t>();
    J.delta = dl.as<int64_t>();
    std::vector<PmJob *> jobs{&J};
    if ((s = pm_basis_batch(p, jobs, st, nullptr, nullptr)) != RR_OK) return s;
    return rank_from_device_basis(M.as<uint32_t>(), dl.as<int64_t>(), b, d + 1, p, sc.as<uint32_t>(), st, rank);
}

// ---------------------------------------------------------------------------

rr_status TickRunner::init(hipStream_t st_, int cus_) {
    st = st_;
    cus = cus_;
    RR_HIP(hipHostMalloc((void **)&stage_host, STAGE_BYTES * NSTAGE, hipHostMallocDefault));
    for (int i = 0; i < NSTAGE; i++) RR_HIP(hipEventCreateWithFlags(&stage_ev[i], hipEventDisableTiming));
    rr_status s;
    if ((s = tiles_dev.alloc(STAGE_BYTES)) != RR_OK) return s;
    return RR_OK;
}

TickRunner::~TickRunner() {
    if (stage_host) {
        (void)hipStreamSynchronize(st);
        (void)hipHostFree(stage_host);
    }
    for (int i = 0; i < NSTAGE; i++)
        if (stage_ev[i]) (void)hipEventDestroy(stage_ev[i]);
}

template <int B>
static void launch_tiles(const Tile *tiles, uint32_t n, uint32_t grid, hipStream_t st, const ModP &mp) {
    hipLaunchKernelGGL(k_tiles<B>, dim3(grid), dim3(WG), 0, st, tiles, n, mp.p, mp.pd, mp.inv);
}

// One tick: the next entry of every request in `reqs` (all in the Krylov stage).
rr_status TickRunner::tick(std::vector<Request *> &reqs, uint32_t p) {
    if (reqs.empty()) return RR_OK;
    ModP mp = ModP::make(p);
    // group tiles by lane width; assign partial slots
    std::vector<Tile> by_b[7];
    std::vector<RedJob> jobs;
    size_t part_words = 0;
    for (auto *R : reqs) {
        const auto &src = R->next_tiles();
        part_words += src.size() * R->P;
    }
    rr_status s;
    if (part_words > part_cap) {
        RR_HIP(hipStreamSynchronize(st));
        part.~DevBuf();
        new (&part) DevBuf();
        part_cap = std::max(part_words, part_cap * 2);
        if ((s = part.alloc(part_cap * 4)) != RR_OK) return s;
    }
    size_t off = 0;
    for (auto *R : reqs) {
        const auto &src = R->next_tiles();
        int lb = 0;
        while ((1u << lb) < R->B) lb++;
        RedJob J;
        J.part = part.as<uint32_t>() + off;
        J.ntiles = (uint32_t)src.size();
        J.P = R->P;
        J.out = R->seq.as<uint32_t>() + R->t * R->P;
        for (const Tile &T0 : src) {
            Tile T = T0;
            T.part = part.as<uint32_t>() + off;
            off += R->P;
            by_b[lb].push_back(T);
        }
        jobs.push_back(J);
    }
    // stage descriptors: [tiles by width][jobs] in one pinned block, one copy
    size_t ntiles_total = 0;
    for (auto &v : by_b) ntiles_total += v.size();
    size_t bytes = ntiles_total * sizeof(Tile) + jobs.size() * sizeof(RedJob);
    RR_CHECK(bytes <= STAGE_BYTES, RR_ENOMEM, "tick descriptors exceed staging (%zu bytes)", bytes);
    int si = stage_next;
    stage_next = (stage_next + 1) % NSTAGE;
    RR_HIP(hipEventSynchronize(stage_ev[si]));  // the copy that last used this slot is done
    char *h = stage_host + (size_t)si * STAGE_BYTES;
    size_t pos = 0;
    size_t tile_off[7];
    for (int lb = 0; lb < 7; lb++) {
        tile_off[lb] = pos;
        if (!by_b[lb].empty()) std::memcpy(h + pos, by_b[lb].data(), by_b[lb].size() * sizeof(Tile));
        pos += by_b[lb].size() * sizeof(Tile);
    }
    size_t job_off = pos;
    std::memcpy(h + pos, jobs.data(), jobs.size() * sizeof(RedJob));
    RR_HIP(hipMemcpyAsync(tiles_dev.ptr, h, bytes, hipMemcpyHostToDevice, st));
    RR_HIP(hipEventRecord(stage_ev[si], st));
    const char *dbase = tiles_dev.as<char>();
    for (int lb = 0; lb < 7; lb++) {
        uint32_t n = (uint32_t)by_b[lb].size();
        if (!n) continue;
        const Tile *tl = (const Tile *)(dbase + tile_off[lb]);
        uint32_t grid = std::min<uint32_t>(n, (uint32_t)cus * 4);
        switch (lb) {
            case 0: launch_tiles<1>(tl, n, grid, st, mp); break;
            case 1: launch_tiles<2>(tl, n, grid, st, mp); break;
            case 2: launch_tiles<4>(tl, n, grid, st, mp); break;
            case 3: launch_tiles<8>(tl, n, grid, st, mp); break;
            case 4: launch_tiles<16>(tl, n, grid, st, mp); break;
            case 5: launch_tiles<32>(tl, n, grid, st, mp); break;
            default: launch_tiles<64>(tl, n, grid, st, mp); break;
        }
        RR_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_reduce_jobs, dim3((uint32_t)jobs.size()), dim3(256), 0, st, (const RedJob *)(dbase + job_off),
                       p);
    RR_HIP(hipGetLastError());
    for (auto *R : reqs) R->t++;
    ticks++;
    return RR_OK;
}

}  // namespace rr

using namespace rr;

extern "C" {

rr_status rr_batch_rank(uint32_t count, const rr_csr *mats, uint32_t p, const uint32_t *b, const uint32_t *const *dr,
                        const uint32_t *const *dc, const uint32_t *const *V, const uint64_t *seq_len,
                        const uint64_t *order, int64_t *rank_out, rr_batch_stats *stats) {
    clear_error();
    RR_CHECK(count >= 1 && mats && b && dr && dc && V && rank_out, RR_EINVAL, "null argument");
    auto t0 = std::chrono::steady_clock::now();
    int dev = 0, cus = 0;
    RR_HIP(hipGetDevice(&dev));
    RR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    hipStream_t st;
    RR_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct SG {
        hipStream_t s;
        ~SG() { (void)hipStreamDestroy(s); }
    } guard{st};
    std::vector<std::unique_ptr<Request>> reqs(count);
    rr_status s;
    for (uint32_t i = 0; i < count; i++) {
        reqs[i].reset(new Request());
        uint64_t L = seq_len && seq_len[i] ? seq_len[i] : 2 * std::min(mats[i].m, mats[i].n) / b[i] + 8;
        uint64_t d = order && order[i] ? order[i] : L - 1;
        if ((s = reqs[i]->init(&mats[i], p, b[i], dr[i], dc[i], V[i], L, d, 2048, st)) != RR_OK) {
            set_error("matrix %u: %s", i, rr_last_error());
            return s;
        }
    }
    RR_HIP(hipStreamSynchronize(st));
    auto t1 = std::chrono::steady_clock::now();
    // Krylov: ticks until every sequence is complete
    TickRunner T;
    if ((s = T.init(st, cus)) != RR_OK) return s;
    std::vector<Request *> active;
    for (;;) {
        active.clear();
        for (auto &r : reqs)
            if (r->stage == STAGE_KRYLOV && r->t < r->L) active.push_back(r.get());
        if (active.empty()) break;
        if ((s = T.tick(active, p)) != RR_OK) return s;
    }
    RR_HIP(hipStreamSynchronize(st));
    auto t2 = std::chrono::steady_clock::now();
    // PM-basis of every request, batched rounds
    std::vector<PmJob *> jobs;
    for (auto &r : reqs) {
        if ((s = r->start_pm(st)) != RR_OK) return s;
        jobs.push_back(&r->pm);
    }
    uint64_t rounds = 0, launches = 0;
    if ((s = pm_basis_batch(p, jobs, st, &rounds, &launches)) != RR_OK) return s;
    for (uint32_t i = 0; i < count; i++) {
        if ((s = reqs[i]->finish_rank(st)) != RR_OK) return s;
        rank_out[i] = reqs[i]->rank;
    }
    auto t3 = std::chrono::steady_clock::now();
    if (stats) {
        auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
        stats->ms_setup = ms(t0, t1);
        stats->ms_krylov = ms(t1, t2);
        stats->ms_pm = ms(t2, t3);
        stats->ms_total = ms(t0, t3);
        stats->ticks = T.ticks;
        stats->pm_rounds = rounds;
        stats->pm_launches = launches;
    }
    return RR_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Autotuner: block width b (and with it rows per wavefront, 64/b) per matrix.
// Measured once per process on the device:
//   c_B   Krylov pass time per nonzero for lane width B (tile kernel on a
//         probe matrix with +-1/+-2 entries, 16 ticks timed with events),
//   pm_b  PM-basis time per order d = 255 for block width b.
// Predicted device time of matrix i with width b:
//   L(b) * nnz_i * c_b + pm_b * (d(b) / 255) * log2(d(b) + 1) / 8,
//   L(b) = 2 min(m, n) / b + 8, d(b) = L(b) - 1.

namespace rr {

struct TuneProbe {
    bool done = false;
    double c[3] = {0, 0, 0}, pm[3] = {0, 0, 0};
};
static TuneProbe g_probe;
static const uint32_t TUNE_B[3] = {8, 16, 32};

static rr_status run_probe(uint32_t p, int cus, hipStream_t st) {
    const uint64_t m = 24000, n = 21600, k = 11;
    std::vector<uint64_t> rp(m + 1);
    std::vector<uint32_t> col(m * k), val(m * k);
    uint64_t x = 0x9E3779B97F4A7C15ull;
    auto rnd = [&]() {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        return x;
    };
    for (uint64_t r = 0; r < m; r++) {
        rp[r] = r * k;
        for (uint64_t e = 0; e < k; e++) {
            col[r * k + e] = (uint32_t)(rnd() % n);
            uint32_t mag = (rnd() & 3) == 0 ? 2 : 1;
            val[r * k + e] = (rnd() & 1) ? p - mag : mag;
        }
    }
    rp[m] = m * k;
    std::vector<uint32_t> dr(m, 3), dc(n, 5), V(n * 32);
    for (auto &v : V) v = (uint32_t)(rnd() % p);
    rr_csr A{m, n, rp.data(), col.data(), val.data()};
    hipEvent_t e0, e1;
    RR_HIP(hipEventCreate(&e0));
    RR_HIP(hipEventCreate(&e1));
    rr_status s;
    for (int i = 0; i < 3; i++) {
        uint32_t b = TUNE_B[i];
        std::vector<uint32_t> Vb(n * b);
        for (uint64_t r = 0; r < n; r++) std::memcpy(&Vb[r * b], &V[r * 32], b * 4);
        Request R;
        if ((s = R.init(&A, p, b, dr.data(), dc.data(), Vb.data(), 40, 39, 2048, st)) != RR_OK) return s;
        TickRunner T;
        if ((s = T.init(st, cus)) != RR_OK) return s;
        std::vector<Request *> one{&R};
        for (int w = 0; w < 4; w++)
            if ((s = T.tick(one, p)) != RR_OK) return s;
        RR_HIP(hipEventRecord(e0, st));
        for (int w = 0; w < 16; w++)
            if ((s = T.tick(one, p)) != RR_OK) return s;
        RR_HIP(hipEventRecord(e1, st));
        RR_HIP(hipEventSynchronize(e1));
        float ms = 0;
        RR_HIP(hipEventElapsedTime(&ms, e0, e1));
        g_probe.c[i] = (ms / 16.0) / (double)(m * k);
        // PM-basis probe, order 255, random sequence
        const uint64_t d = 255, P = b * (b + 1) / 2;
        std::vector<uint32_t> seq((d + 1) * P);
        for (auto &v : seq) v = (uint32_t)(rnd() % p);
        DevBuf dseq, G, M, dl;
        if ((s = dseq.alloc(seq.size() * 4)) != RR_OK) return s;
        if ((s = G.alloc((size_t)2 * b * b * (d + 1) * 4)) != RR_OK) return s;
        if ((s = M.alloc((size_t)4 * b * b * (d + 1) * 4)) != RR_OK) return s;
        std::vector<int64_t> delta(2 * b);
        RR_HIP(hipMemcpyAsync(dseq.ptr, seq.data(), seq.size() * 4, hipMemcpyHostToDevice, st));
        if ((s = build_G_device(dseq.as<uint32_t>(), b, d + 1, G.as<uint32_t>(), st)) != RR_OK) return s;
        RR_HIP(hipStreamSynchronize(st));
        auto t0 = std::chrono::steady_clock::now();
        if ((s = pm_basis_device(p, b, d, G.as<uint32_t>(), M.as<uint32_t>(), delta.data(), st)) != RR_OK) return s;
        g_probe.pm[i] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    g_probe.done = true;
    return RR_OK;
}

}  // namespace rr

extern "C" {

rr_status rr_autotune_b(uint32_t count, const rr_csr *mats, uint32_t p, uint32_t *b_out, double *pred_ms) {
    clear_error();
    RR_CHECK(count == 0 || (mats && b_out), RR_EINVAL, "null argument");
    RR_CHECK(p == NTT_PRIME, RR_EINVAL, "the autotuner's PM-basis probe needs p = 2013265921");
    if (!g_probe.done) {
        int dev = 0, cus = 0;
        RR_HIP(hipGetDevice(&dev));
        RR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        hipStream_t st;
        RR_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        rr_status s = run_probe(p, cus, st);
        (void)hipStreamDestroy(st);
        if (s != RR_OK) return s;
    }
    for (uint32_t i = 0; i < count; i++) {
        const rr_csr &A = mats[i];
        uint64_t nnz = A.row_ptr[A.m], mn = std::min(A.m, A.n);
        double best = 1e300;
        uint32_t bb = TUNE_B[0];
        for (int k = 0; k < 3; k++) {
            uint32_t b = TUNE_B[k];
            double L = 2.0 * (double)mn / b + 8, d = L - 1;
            if (d + 1 > (double)(1u << NTT_MAX_LOGN)) continue;
            double t = L * (double)nnz * g_probe.c[k] + g_probe.pm[k] * (d / 255.0) * std::log2(d + 1) / 8.0;
            if (t < best) {
                best = t;
                bb = b;
            }
        }
        b_out[i] = bb;
        if (pred_ms) pred_ms[i] = best;
    }
    return RR_OK;
}

rr_status rr_autotune_probe(double *c_ms_per_nnz, double *pm_ms_d255) {
    RR_CHECK(c_ms_per_nnz && pm_ms_d255, RR_EINVAL, "null argument");
    RR_CHECK(g_probe.done, RR_ESTATE, "autotuner has not probed yet");
    for (int k = 0; k < 3; k++) {
        c_ms_per_nnz[k] = g_probe.c[k];
        pm_ms_d255[k] = g_probe.pm[k];
    }
    return RR_OK;
}

}  // extern "C"
