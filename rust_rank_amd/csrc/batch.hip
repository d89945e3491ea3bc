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
#include <thread>
#include <atomic>
#include <string>

namespace rr {

// ---------------------------------------------------------------------------
// kernels

// t = 0 projections x^T x of row ranges of several requests' vectors.
template <int B>
__global__ __launch_bounds__(WG) void k_gram_tiles(const GramTile *__restrict__ tiles, uint32_t ntiles, uint32_t p,
                                                   double pd, double inv) {
    constexpr int G = 64 / B;
    __shared__ uint32_t lds[Gram<B>::LDS_WORDS];
    __shared__ WaveLds wl[WG / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int slot = lane / B, j = lane & (B - 1);
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const GramTile T = tiles[t];
        Gram<B> g;
        g.init();
        for (uint32_t base = T.row_beg + wave * G; base < T.row_end; base += (WG / 64) * G) {
            const uint32_t r = base + slot;
            const uint32_t o = r < T.row_end ? as_global(T.in)[(size_t)r * B + j] : 0u;
            g.add(o, o, lane, wl[wave], pd, inv);
        }
        g.finish(T.part, T.b, p, pd, inv, lds);
        __syncthreads();  // lds is reused by the next tile
    }
}

// out[ii] = sum_k part[k * P + ii] mod p.  Workgroup (job, y) sums columns
// [64 y, 64 y + 64) of its job: lanes over the columns, its 16 wavefronts over
// partials k = w mod 16 with four loads in flight per lane, the sums
// combined in LDS.
constexpr uint32_t RJ_WAVES = 16;
__global__ __launch_bounds__(RJ_WAVES * 64) void k_reduce_jobs(const RedJob *__restrict__ jobs, uint32_t p) {
    __shared__ uint64_t red[RJ_WAVES][64];
    const RedJob J = jobs[blockIdx.x];
    const uint32_t c0 = blockIdx.y * 64;
    if (c0 >= J.P) return;
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t ii = c0 + lane, ic = min(ii, J.P - 1);
    const GLOBAL_AS uint32_t *src = as_global(J.part) + ic;
    uint64_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;  // partials < 2^31, fewer than 2^32 of them
    uint32_t k = w;
    for (; k + 3 * RJ_WAVES < J.ntiles; k += 4 * RJ_WAVES) {
        s0 += src[(size_t)k * J.P];
        s1 += src[(size_t)(k + RJ_WAVES) * J.P];
        s2 += src[(size_t)(k + 2 * RJ_WAVES) * J.P];
        s3 += src[(size_t)(k + 3 * RJ_WAVES) * J.P];
    }
    for (; k < J.ntiles; k += RJ_WAVES) s0 += src[(size_t)k * J.P];
    red[w][lane] = s0 + s1 + s2 + s3;
    __syncthreads();
    if (w == 0 && ii < J.P) {
        uint64_t t = 0;
#pragma unroll
        for (uint32_t q = 0; q < RJ_WAVES; q++) t += red[q][lane];
        as_global(J.out)[ii] = (uint32_t)(t % p);
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
    uint64_t words = 3 * nnz + 2 * (m + n) + (packed ? 0 : 3 * nnz) + m + 2 * n  // sliced operands + scales
                     + B * (2 * n + m)                                                   // z, x0, y
                     + L * P                                                             // sequence
                     + (uint64_t)2 * b * b * (d + 1) + (uint64_t)4 * b * b * (d + 1)     // G, M
                     + (uint64_t)2 * (4ull * b * b * (d / 2 + 2) + 2ull * b * b * (d / 2 + 2)) + 2 * b * b + 64
                     + (uint64_t)2 * 4ull * b * b * 2 * (d + 1);  // kept spectra along the recursion
    return words * 4;
}

uint64_t Request::device_bytes() const { return request_bytes(m, n, nnz, b, L, d, packed); }

rr_status Request::init(const rr_csr *A, uint32_t p_, uint32_t b_, const uint32_t *dr, const uint32_t *dcv,
                        const uint32_t *V, uint64_t L_, uint64_t d_, uint32_t tile_rows, hipStream_t st,
                        uint32_t su_req, uint32_t steps_per_wave) {
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
    RR_CHECK(su_req == 0 || su_req == 3 || su_req == 4, RR_EINVAL, "su=%u must be 3 or 4 (0: chosen)", su_req);
    RR_CHECK(m >= 1 && n >= 1 && m < (1ull << 31) && n < (1ull << 31), RR_EINVAL, "matrix size out of range");
    RR_CHECK(L >= 2 && d >= 1 && d < L, RR_EINVAL, "need 1 <= d < L (L=%llu d=%llu)", (unsigned long long)L,
             (unsigned long long)d);
    RR_CHECK(d + 1 <= (1ull << NTT_MAX_LOGN), RR_EINVAL, "order above the transform limit");
    RR_CHECK(A->row_ptr && A->row_ptr[0] == 0, RR_EINVAL, "row_ptr[0] must be 0");
    for (uint64_t r = 0; r < m; r++) RR_CHECK(A->row_ptr[r + 1] >= A->row_ptr[r], RR_EINVAL, "row_ptr not monotone");
    nnz = A->row_ptr[m];
    RR_CHECK(nnz < (1ull << 32), RR_EINVAL, "nnz must be < 2^32");
    RR_CHECK(std::max(m, n) * B * 4 < 0xFFFFFF00ull, RR_EINVAL, "block vectors reach 4 GiB");
    RR_CHECK(nnz == 0 || (A->col && A->val), RR_EINVAL, "null col/val");
    for (uint64_t i = 0; i < m; i++) RR_CHECK(dr[i] >= 1 && dr[i] < p, RR_EINVAL, "row_precond out of range");
    for (uint64_t i = 0; i < n; i++) RR_CHECK(dcv[i] >= 1 && dcv[i] < p, RR_EINVAL, "col_precond out of range");

    std::vector<uint32_t> dc2(n);
    {
        const HostMod hm(p);
        for (uint64_t i = 0; i < n; i++) dc2[i] = hm.mul(dcv[i], dcv[i]);
    }
    rr_status s;
    if ((s = upload_vec(dr_d, std::vector<uint32_t>(dr, dr + m), st)) != RR_OK) return s;
    if ((s = upload_vec(dc_d, std::vector<uint32_t>(dcv, dcv + n), st)) != RR_OK) return s;
    if ((s = upload_vec(dc2_d, dc2, st)) != RR_OK) return s;
    // V < p is checked on the device as it is uploaded (padded to stride B there)
    if ((s = upload_block_vector(x0, V, n, b, B, p, "V", st)) != RR_OK) return s;
    if ((s = z.alloc((size_t)n * B * 4)) != RR_OK) return s;
    if ((s = y.alloc((size_t)m * B * 4)) != RR_OK) return s;
    if ((s = seq.alloc((size_t)L * P * 4)) != RR_OK) return s;
    RR_HIP(hipMemsetAsync(y.ptr, 0, (size_t)m * B * 4, st));
    size_t tot = (size_t)n * B;
    ModP mp = ModP::make(p);
    hipLaunchKernelGGL(k_scale_rows_b, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, x0.as<uint32_t>(),
                       dc_d.as<uint32_t>(), z.as<uint32_t>(), (uint32_t)n, B, p, mp.inv);
    RR_HIP(hipGetLastError());

    // Sliced operands for nwg workgroups (a multiple of the XCD count, so a
    // job spread over the XCDs keeps its XCD placement), checked and laid
    // out on the device.
    const uint32_t RPW = 64 / (B / (B < 4 ? B : 4));
    const uint64_t steps = std::max(nnz, std::max(m, n)) / RPW + 1;
    RR_CHECK(steps_per_wave >= 1, RR_EINVAL, "steps per wavefront must be positive");
    const uint64_t waves = (steps + steps_per_wave - 1) / steps_per_wave;
    int dev = 0, cus = 0;
    RR_HIP(hipGetDevice(&dev));
    RR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    OperandInfo oi;
    s = setup_operands(
        A, p, B, su_req,
        [&](bool pk, uint32_t su_) {
            const uint64_t max_wg = (uint64_t)cus * spmm_blocks_per_cu(B, pk, su_);
            nwg = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(NXCD, (waves / (WG / 64) + NXCD - 1) / NXCD * NXCD),
                                               std::max<uint64_t>(NXCD, max_wg / NXCD * NXCD));
            return nwg * (uint32_t)(WG / 64);
        },
        dr, dc2.data(), sa, sat, &oi, st);
    if (s != RR_OK) return s;
    packed = oi.packed;
    su = oi.su;
    args[0] = spmm_args(sa, z.as<uint32_t>(), n, B, y.as<uint32_t>(), dr_d.as<uint32_t>(), p, b);
    args[1] = spmm_args(sat, y.as<uint32_t>(), m, B, z.as<uint32_t>(), dc2_d.as<uint32_t>(), p, b);

    // t = 0 projection tiles
    tiles_gram.clear();
    for (uint64_t r0 = 0; r0 < n; r0 += tile_rows) {
        GramTile T;
        T.in = x0.as<uint32_t>();
        T.part = nullptr;
        T.row_beg = (uint32_t)r0;
        T.row_end = (uint32_t)std::min<uint64_t>(n, r0 + tile_rows);
        T.b = b;
        T.pad = 0;
        tiles_gram.push_back(T);
    }
    t = 0;
    stage = STAGE_KRYLOV;
    return RR_OK;
}

rr_status Request::alloc_pm() {
    rr_status s;
    if ((s = G.alloc((size_t)2 * b * b * (d + 1) * 4)) != RR_OK) return s;
    if ((s = M.alloc((size_t)4 * b * b * (d + 1) * 4)) != RR_OK) return s;
    if ((s = delta.alloc((size_t)2 * b * 8)) != RR_OK) return s;
    if ((s = rank_in.alloc((size_t)(2 * b * b + 1) * 4)) != RR_OK) return s;
    pm.b = b;
    pm.d = d;
    pm.G = G.as<uint32_t>();
    pm.M = M.as<uint32_t>();
    pm.delta = delta.as<int64_t>();
    pm.avail = 0;
    return RR_OK;
}

rr_status Request::grow_G(hipStream_t st) {
    const uint64_t k1 = std::min<uint64_t>(t, d + 1);
    if (pm.avail == ~0ull || k1 <= pm.avail) return RR_OK;
    rr_status s = build_G_range(seq.as<uint32_t>(), b, d + 1, pm.avail, k1, G.as<uint32_t>(), st);
    if (s != RR_OK) return s;
    pm.avail = k1 == d + 1 ? ~0ull : k1;
    return RR_OK;
}

rr_status Request::start_pm(hipStream_t st, bool release) {
    rr_status s;
    RR_CHECK(t > d, RR_EINVAL, "sequence of %llu entries too short for order %llu", (unsigned long long)t,
             (unsigned long long)d);
    if (!G.ptr && (s = alloc_pm()) != RR_OK) return s;
    if ((s = grow_G(st)) != RR_OK) return s;
    stage = STAGE_PM;
    if (!release) return RR_OK;
    // the operands are no longer needed; the device memory cache hands their
    // blocks out again only once the passes queued before this have run
    sa.~SliceMat();
    new (&sa) SliceMat();
    sat.~SliceMat();
    new (&sat) SliceMat();
    return RR_OK;
}

// Rank from a device-resident basis [2b][2b][len] and delta: the two coefficient
// matrices the formula needs are extracted on the device (2b^2 + 1 words come back).
rr_status rank_extract_async(const uint32_t *M, const int64_t *delta, uint32_t b, uint64_t len, uint32_t *scratch,
                             uint32_t *host, hipStream_t st) {
    hipLaunchKernelGGL(k_rank_extract, dim3(1), dim3(256), 0, st, M, delta, b, len, scratch);
    RR_HIP(hipGetLastError());
    RR_HIP(hipMemcpyAsync(host, scratch, ((size_t)2 * b * b + 1) * 4, hipMemcpyDeviceToHost, st));
    return RR_OK;
}

int64_t rank_from_extracted(const uint32_t *h, uint32_t b, uint32_t p) {
    const uint64_t K = h[0];
    std::vector<uint32_t> F0(h + 1, h + 1 + b * b), FK(h + 1 + b * b, h + 1 + 2 * b * b);
    const uint64_t r = (uint64_t)b * (K - 3) + host_rank_bxb(FK, b, p) + host_rank_bxb(F0, b, p);
    return (int64_t)r;  // the reference's usize arithmetic wraps the same way when K < 3
}

rr_status rank_from_device_basis(const uint32_t *M, const int64_t *delta, uint32_t b, uint64_t len, uint32_t p,
                                 uint32_t *scratch, hipStream_t st, int64_t *rank) {
    std::vector<uint32_t> h(2 * b * b + 1);
    rr_status s = rank_extract_async(M, delta, b, len, scratch, h.data(), st);
    if (s != RR_OK) return s;
    RR_HIP(hipStreamSynchronize(st));
    *rank = rank_from_extracted(h.data(), b, p);
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
    if ((s = desc_dev.alloc(STAGE_BYTES)) != RR_OK) return s;
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
static void launch_gram_tiles(const GramTile *tiles, uint32_t n, uint32_t grid, hipStream_t st, const ModP &mp) {
    hipLaunchKernelGGL(k_gram_tiles<B>, dim3(grid), dim3(WG), 0, st, tiles, n, mp.p, mp.pd, mp.inv);
}

// nticks ticks: the next nticks entries of every request in `reqs` (all in
// the Krylov stage).  Requests at t = 0 contribute projection tiles (then
// nticks must be 1); the others one pass job per tick, grouped by (lane
// width, value kind, block size) into one k_spmm launch per group and tick.
// The descriptors of all the ticks go down in one copy, and one
// k_reduce_jobs launch folds every request's partials of every tick (a
// request's partials are always reduced alone).
rr_status TickRunner::tick(std::vector<Request *> &reqs, uint32_t p, uint32_t nticks) {
    launched = false;
    if (reqs.empty() || nticks == 0) return RR_OK;
    for (auto *R : reqs)
        RR_CHECK(!fail_id || R->id != fail_id, RR_ENOMEM, "injected tick failure (request %llu)",
                 (unsigned long long)R->id);
    // Requests whose descriptors do not fit one staging slot together run in
    // several staged launches (a request's results do not depend on which
    // requests share its launches).
    auto bytes_of = [&](const Request *R) {
        return R->t == 0 ? R->tiles_gram.size() * sizeof(GramTile) + sizeof(RedJob) + 64
                         : (size_t)nticks * (sizeof(SpmmArgs) + sizeof(RedJob) + 64) + NXCD * 16 + 64;
    };
    const size_t cap = STAGE_BYTES - 7 * 16 - 1024 - 2 * 16 * (WGMAP_HEAD_WORDS * 4 + 16);
    size_t tot = 0;
    for (auto *R : reqs) tot += bytes_of(R);
    if (tot <= cap) return tick_staged(reqs, p, nticks);
    std::vector<Request *> part_reqs;
    size_t acc = 0;
    for (size_t i = 0; i <= reqs.size(); i++) {
        if (i == reqs.size() || (acc + bytes_of(reqs[i]) > cap && !part_reqs.empty())) {
            const bool before = launched;
            rr_status s = tick_staged(part_reqs, p, nticks);
            launched |= before;
            if (s != RR_OK) return s;
            part_reqs.clear();
            acc = 0;
            if (i == reqs.size()) break;
        }
        RR_CHECK(bytes_of(reqs[i]) <= cap, RR_ENOMEM, "one request's tick descriptors exceed staging");
        part_reqs.push_back(reqs[i]);
        acc += bytes_of(reqs[i]);
    }
    return RR_OK;
}

rr_status TickRunner::tick_staged(std::vector<Request *> &reqs, uint32_t p, uint32_t nticks) {
    if (reqs.empty() || nticks == 0) return RR_OK;
    ModP mp = ModP::make(p);
    for (auto *R : reqs)
        RR_CHECK(R->t + nticks <= R->L && (nticks == 1 || R->t > 0), RR_EINVAL, "tick chunk past a sequence");
    struct Group {
        uint32_t B = 0, su = 0;
        bool packed = false;
        std::vector<SpmmArgs> jobs;
        std::vector<uint32_t> nwg;  // workgroups of each job
    };
    std::vector<std::vector<Group>> groups(nticks);
    std::vector<GramTile> gram_by_b[7];
    std::vector<RedJob> red;
    size_t part_words = 0;
    for (auto *R : reqs)
        part_words += (size_t)(R->t == 0 ? R->tiles_gram.size() : (size_t)R->nwg * nticks) * R->P;
    rr_status s;
    if (part_words > part_cap) {
        RR_HIP(hipStreamSynchronize(st));
        part.~DevBuf();
        new (&part) DevBuf();
        part_cap = std::max(part_words, part_cap * 2);
        if ((s = part.alloc(part_cap * 4)) != RR_OK) return s;
    }
    uint32_t *pbase = part.as<uint32_t>();
    size_t off = 0;
    for (uint32_t c = 0; c < nticks; c++) {
        for (auto *R : reqs) {
            const uint64_t t = R->t + c;
            RedJob J;
            J.part = pbase + off;
            J.P = R->P;
            J.out = R->seq.as<uint32_t>() + t * R->P;
            if (t == 0) {
                int lb = 0;
                while ((1u << lb) < R->B) lb++;
                J.ntiles = (uint32_t)R->tiles_gram.size();
                for (const GramTile &T0 : R->tiles_gram) {
                    GramTile T = T0;
                    T.part = pbase + off;
                    off += R->P;
                    gram_by_b[lb].push_back(T);
                }
            } else {
                Group *G = nullptr;
                for (auto &g : groups[c])
                    if (g.B == R->B && g.su == R->su && g.packed == R->packed) G = &g;
                if (!G) {
                    groups[c].emplace_back();
                    G = &groups[c].back();
                    G->B = R->B;
                    G->su = R->su;
                    G->packed = R->packed;
                }
                SpmmArgs a = R->args[(t & 1) ? 0 : 1];
                a.gram_off = off;
                G->jobs.push_back(a);
                G->nwg.push_back(R->nwg);
                J.ntiles = R->nwg;
                off += (size_t)R->nwg * R->P;
            }
            red.push_back(J);
        }
    }
    // Every tick has the same groups with the same jobs' workgroups (the same
    // requests in the same order), so one workgroup map per group serves all.
    std::vector<std::vector<uint32_t>> maps(groups[0].size());
    std::vector<uint32_t> grids(groups[0].size());
    for (size_t gi = 0; gi < groups[0].size(); gi++) maps[gi] = build_wg_map(groups[0][gi].nwg, &grids[gi]);
    for (uint32_t c = 1; c < nticks; c++) {
        RR_CHECK(groups[c].size() == groups[0].size(), RR_EINVAL, "tick groups differ");
        for (size_t gi = 0; gi < groups[c].size(); gi++)
            RR_CHECK(groups[c][gi].nwg == groups[0][gi].nwg, RR_EINVAL, "tick groups differ");
    }
    // stage descriptors in one pinned block, one copy:
    // [gram tiles by width][workgroup map per group][per tick, per group: jobs][reduction jobs]
    auto align = [](size_t x) { return (x + 15) & ~(size_t)15; };
    size_t bytes = 0;
    size_t gram_off[7];
    std::vector<std::vector<size_t>> job_off(nticks);
    std::vector<size_t> map_off;
    for (int lb = 0; lb < 7; lb++) {
        gram_off[lb] = bytes;
        bytes = align(bytes + gram_by_b[lb].size() * sizeof(GramTile));
    }
    for (const auto &mw : maps) {
        map_off.push_back(bytes);
        bytes = align(bytes + mw.size() * sizeof(uint32_t));
    }
    for (uint32_t c = 0; c < nticks; c++)
        for (size_t gi = 0; gi < groups[c].size(); gi++) {
            job_off[c].push_back(bytes);
            bytes = align(bytes + groups[c][gi].jobs.size() * sizeof(SpmmArgs));
        }
    const size_t red_off = bytes;
    bytes += red.size() * sizeof(RedJob);
    RR_CHECK(bytes <= STAGE_BYTES, RR_ENOMEM, "tick descriptors exceed staging (%zu bytes)", bytes);
    int si = stage_next;
    stage_next = (stage_next + 1) % NSTAGE;
    RR_HIP(hipEventSynchronize(stage_ev[si]));  // the copy that last used this slot is done
    char *h = stage_host + (size_t)si * STAGE_BYTES;
    for (int lb = 0; lb < 7; lb++)
        if (!gram_by_b[lb].empty())
            std::memcpy(h + gram_off[lb], gram_by_b[lb].data(), gram_by_b[lb].size() * sizeof(GramTile));
    for (size_t gi = 0; gi < maps.size(); gi++)
        std::memcpy(h + map_off[gi], maps[gi].data(), maps[gi].size() * sizeof(uint32_t));
    for (uint32_t c = 0; c < nticks; c++)
        for (size_t gi = 0; gi < groups[c].size(); gi++) {
            const Group &G = groups[c][gi];
            std::memcpy(h + job_off[c][gi], G.jobs.data(), G.jobs.size() * sizeof(SpmmArgs));
        }
    std::memcpy(h + red_off, red.data(), red.size() * sizeof(RedJob));
    launched = true;  // from here on, work of this tick is on the stream
    RR_HIP(hipMemcpyAsync(desc_dev.ptr, h, bytes, hipMemcpyHostToDevice, st));
    RR_HIP(hipEventRecord(stage_ev[si], st));
    const char *dbase = desc_dev.as<char>();
    for (int lb = 0; lb < 7; lb++) {
        uint32_t n = (uint32_t)gram_by_b[lb].size();
        if (!n) continue;
        const GramTile *tl = (const GramTile *)(dbase + gram_off[lb]);
        uint32_t grid = std::min<uint32_t>(n, (uint32_t)cus * 4);
        switch (lb) {
            case 0: launch_gram_tiles<1>(tl, n, grid, st, mp); break;
            case 1: launch_gram_tiles<2>(tl, n, grid, st, mp); break;
            case 2: launch_gram_tiles<4>(tl, n, grid, st, mp); break;
            case 3: launch_gram_tiles<8>(tl, n, grid, st, mp); break;
            case 4: launch_gram_tiles<16>(tl, n, grid, st, mp); break;
            case 5: launch_gram_tiles<32>(tl, n, grid, st, mp); break;
            default: launch_gram_tiles<64>(tl, n, grid, st, mp); break;
        }
        RR_HIP(hipGetLastError());
    }
    for (uint32_t c = 0; c < nticks; c++)
        for (size_t gi = 0; gi < groups[c].size(); gi++) {
            const Group &G = groups[c][gi];
            launch_spmm(G.B, G.packed, G.su, (const SpmmArgs *)(dbase + job_off[c][gi]),
                        (const uint32_t *)(dbase + map_off[gi]), grids[gi], pbase, st);
            RR_HIP(hipGetLastError());
        }
    uint32_t maxP = 0;
    for (const RedJob &J : red) maxP = std::max(maxP, J.P);
    RR_CHECK(red.size() < 65536 * 32768ull, RR_EINVAL, "too many reductions");
    hipLaunchKernelGGL(k_reduce_jobs, dim3((uint32_t)red.size(), (maxP + 63) / 64), dim3(RJ_WAVES * 64), 0, st,
                       (const RedJob *)(dbase + red_off), p);
    RR_HIP(hipGetLastError());
    for (auto *R : reqs) R->t += nticks;
    ticks += nticks;
    return RR_OK;
}

// Entries per rr_krylov_run call and PM-basis rounds kept queued by rank_streamed.
constexpr uint64_t STREAM_CHUNK = 128;
constexpr size_t STREAM_ROUNDS_AHEAD = 96;

rr_status rank_streamed(rr_krylov *k, uint32_t b, uint64_t L, uint64_t d, uint32_t p, uint32_t *seq_out,
                        int64_t *rank) {
    RR_CHECK(d >= 1 && d < L, RR_EINVAL, "order must satisfy 1 <= order < seq_len");
    const uint64_t P = (uint64_t)b * (b + 1) / 2;
    const hipStream_t sk = krylov_stream(k);
    DevBuf G, M, dl, sc;
    rr_status s;
    if ((s = G.alloc((size_t)2 * b * b * (d + 1) * 4)) != RR_OK) return s;
    if ((s = M.alloc((size_t)4 * b * b * (d + 1) * 4)) != RR_OK) return s;
    if ((s = dl.alloc((size_t)2 * b * 8)) != RR_OK) return s;
    if ((s = sc.alloc((size_t)(2 * b * b + 1) * 4)) != RR_OK) return s;
    hipStream_t sp = nullptr;
    hipEvent_t ev = nullptr;
    struct Guard {
        hipStream_t &sp;
        hipEvent_t &ev;
        ~Guard() {
            if (sp) {
                (void)hipStreamSynchronize(sp);
                rr::stream_destroy(sp);
            }
            if (ev) (void)hipEventDestroy(ev);
        }
    } guard{sp, ev};
    RR_HIP(rr::stream_create(&sp));
    RR_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    uint32_t logn = 2;
    while ((1ull << logn) < d + 1) logn++;
    PmScheduler S;
    if ((s = S.init(p, sp, logn, 2 * b, b)) != RR_OK) return s;
    PmJob J;
    J.b = b;
    J.d = d;
    J.K = leaf_order_streamed(2 * b);
    J.G = G.as<uint32_t>();
    J.M = M.as<uint32_t>();
    J.delta = dl.as<int64_t>();
    if ((s = S.add(&J)) != RR_OK) return s;
    J.avail = 0;
    std::vector<PmJob *> jobs{&J};
    EventRing ring;
    if ((s = ring.init(STREAM_ROUNDS_AHEAD)) != RR_OK) return s;
    uint64_t nrounds = 0, t = 0;
    // RR_PROGRESS=<seconds>: a line on stderr at most that often (long calls)
    const char *pe = getenv("RR_PROGRESS");
    const double every = pe ? atof(pe) : 0.0;
    auto t_last = std::chrono::steady_clock::now();
    auto progress = [&](bool force) {
        if (every <= 0) return;
        const auto now = std::chrono::steady_clock::now();
        if (!force && std::chrono::duration<double>(now - t_last).count() < every) return;
        t_last = now;
        fprintf(stderr, "rr_rank: %llu/%llu sequence entries, %llu PM-basis rounds%s\n", (unsigned long long)t,
                (unsigned long long)L, (unsigned long long)nrounds, J.done ? ", basis done" : "");
        fflush(stderr);
    };
    auto rounds = [&]() -> rr_status {  // the rounds the coefficients allow, within the ring
        while (!J.done && ring.room()) {
            bool launched = false;
            rr_status e = S.round(jobs, &launched);
            if (e != RR_OK) return e;
            if (!launched) break;
            nrounds++;
            if ((e = ring.push(sp)) != RR_OK) return e;
        }
        return RR_OK;
    };
    for (t = 0; t < L;) {
        const uint64_t n = std::min<uint64_t>(STREAM_CHUNK, L - t);
        if ((s = rr_krylov_run(k, n, seq_out + t * P)) != RR_OK) return s;
        const uint64_t k1 = std::min<uint64_t>(t + n, d + 1);
        if (k1 > t) {  // coefficients [t, k1) from this call's entries, then to the PM stream
            if ((s = build_G_range(krylov_device_seq(k), b, d + 1, t, k1, G.as<uint32_t>(), sk, t)) != RR_OK) return s;
            RR_HIP(hipEventRecord(ev, sk));
            RR_HIP(hipStreamWaitEvent(sp, ev, 0));
            J.avail = k1 == d + 1 ? ~0ull : k1;
        }
        t += n;
        if ((s = rounds()) != RR_OK) return s;
        progress(false);
    }
    while (!J.done) {
        if ((s = rounds()) != RR_OK) return s;
        if (!J.done && !ring.room()) std::this_thread::yield();
        if (!J.done && ring.room() && S.waiting) {
            set_error("PM-basis waits for coefficients past the sequence");
            return RR_EINVAL;
        }
        progress(false);
    }
    progress(true);
    if (J.err != RR_OK) return J.err;
    return rank_from_device_basis(M.as<uint32_t>(), dl.as<int64_t>(), b, d + 1, p, sc.as<uint32_t>(), sp, rank);
}

}  // namespace rr

using namespace rr;

extern "C" {

static rr_status batch_rank(uint32_t count, const rr_csr *mats, uint32_t p, const uint32_t *b, const uint32_t *su_of,
                            const uint32_t *const *dr, const uint32_t *const *dc, const uint32_t *const *V,
                            const uint64_t *seq_len, const uint64_t *order, int64_t *rank_out, rr_batch_stats *stats);

rr_status rr_batch_rank(uint32_t count, const rr_csr *mats, uint32_t p, const uint32_t *b, const uint32_t *const *dr,
                        const uint32_t *const *dc, const uint32_t *const *V, const uint64_t *seq_len,
                        const uint64_t *order, int64_t *rank_out, rr_batch_stats *stats) {
    return rr::guarded([&]() -> rr_status {
        clear_error();
        return batch_rank(count, mats, p, b, nullptr, dr, dc, V, seq_len, order, rank_out, stats);
    });
}

rr_status rr_batch_rank_tuned(uint32_t count, const rr_csr *mats, uint32_t p, const rr_tuning *tune,
                              const uint32_t *const *dr, const uint32_t *const *dc, const uint32_t *const *V,
                              const uint64_t *seq_len, const uint64_t *order, int64_t *rank_out,
                              rr_batch_stats *stats) {
    return rr::guarded([&]() -> rr_status {
        clear_error();
        RR_CHECK(count >= 1 && tune, RR_EINVAL, "null argument");
        std::vector<uint32_t> b(count), su(count);
        for (uint32_t i = 0; i < count; i++) {
            b[i] = tune[i].b;
            su[i] = tune[i].su;
        }
        return batch_rank(count, mats, p, b.data(), su.data(), dr, dc, V, seq_len, order, rank_out, stats);
    });
}

static rr_status batch_rank(uint32_t count, const rr_csr *mats, uint32_t p, const uint32_t *b, const uint32_t *su_of,
                            const uint32_t *const *dr, const uint32_t *const *dc, const uint32_t *const *V,
                            const uint64_t *seq_len, const uint64_t *order, int64_t *rank_out, rr_batch_stats *stats) {
    RR_CHECK(count >= 1 && mats && b && dr && dc && V && rank_out, RR_EINVAL, "null argument");
    auto t0 = std::chrono::steady_clock::now();
    int dev = 0, cus = 0;
    RR_HIP(hipGetDevice(&dev));
    RR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    hipStream_t st;
    RR_HIP(rr::stream_create(&st));
    struct SG {
        hipStream_t s;
        ~SG() { rr::stream_destroy(s); }
    } guard{st};
    std::vector<std::unique_ptr<Request>> reqs(count);
    rr_status s;
    auto seq_len_of = [&](uint32_t i) -> uint64_t {
        return seq_len && seq_len[i] ? seq_len[i] : b[i] ? 2 * std::min(mats[i].m, mats[i].n) / b[i] + 8 : 0;
    };
    // Host set-up (validation, transpose, slicing, uploads) runs on worker
    // threads, each with its own stream, while this thread already ticks the
    // matrices that are ready: the largest work (sequence length x nonzeros)
    // is set up first, since its sequence is the longest chain.  A request
    // joins the tick loop once its stream's event says its uploads are done.
    struct Setup {
        std::vector<rr_status> st;
        std::vector<std::string> msg;
        std::vector<hipEvent_t> ev;
        std::unique_ptr<std::atomic<int>[]> ready;  // 0 pending, 1 set up, 2 failed
        std::vector<uint32_t> order;
        std::atomic<uint32_t> next{0};
        std::atomic<bool> stop{false};
        std::vector<hipStream_t> ws;
        std::vector<std::thread> pool;
        ~Setup() {
            stop = true;
            for (auto &t : pool) t.join();
            for (auto w : ws) {
                (void)hipStreamSynchronize(w);
                rr::stream_destroy(w);
            }
            for (auto e : ev)
                if (e) (void)hipEventDestroy(e);
        }
    } su;
    su.st.assign(count, RR_OK);
    su.msg.resize(count);
    su.ev.assign(count, nullptr);
    su.ready.reset(new std::atomic<int>[count]);
    su.order.resize(count);
    for (uint32_t i = 0; i < count; i++) {
        su.ready[i] = 0;
        su.order[i] = i;
        RR_HIP(hipEventCreateWithFlags(&su.ev[i], hipEventDisableTiming));
    }
    std::vector<double> work(count);
    for (uint32_t i = 0; i < count; i++) work[i] = (double)seq_len_of(i) * (double)mats[i].row_ptr[mats[i].m];
    std::stable_sort(su.order.begin(), su.order.end(), [&](uint32_t x, uint32_t y) { return work[x] > work[y]; });
    {
        const uint32_t hw = std::max(2u, std::thread::hardware_concurrency());
        const uint32_t nth = std::max(1u, std::min<uint32_t>(count, std::min<uint32_t>(hw, 16) - 1));
        for (uint32_t w = 0; w < nth; w++) {
            hipStream_t ws;
            RR_HIP(rr::stream_create(&ws));
            su.ws.push_back(ws);
        }
        for (uint32_t w = 0; w < nth; w++)
            su.pool.emplace_back([&, w]() {
                const bool dev_ok = hipSetDevice(dev) == hipSuccess;
                for (uint32_t k; !su.stop && (k = su.next++) < count;) {
                    const uint32_t i = su.order[k];
                    const rr_status e = guarded([&]() -> rr_status {
                        if (!dev_ok) {
                            set_error("set-up thread could not select the device");
                            return RR_EDEVICE;
                        }
                        reqs[i].reset(new Request());
                        const uint64_t L = seq_len_of(i);
                        const uint64_t d = order && order[i] ? order[i] : L - 1;
                        rr_status s = reqs[i]->init(&mats[i], p, b[i], dr[i], dc[i], V[i], L, d, 2048, su.ws[w],
                                                    su_of ? su_of[i] : 0);
                        if (s == RR_OK && hipEventRecord(su.ev[i], su.ws[w]) != hipSuccess) {
                            set_error("event record failed");
                            s = RR_EDEVICE;
                        }
                        return s;
                    });
                    if (e != RR_OK) su.msg[i] = rr_last_error();
                    su.st[i] = e;
                    su.ready[i].store(e == RR_OK ? 1 : 2, std::memory_order_release);
                }
            });
    }
    std::vector<char> joined(count, 0);
    uint32_t njoined = 0;
    auto t1 = t0;
    // Krylov chunks on st and PM-basis rounds on st_pm, issued from this one
    // thread and run side by side on the device: a matrix whose sequence is
    // complete joins the PM scheduler while the longer sequences still tick.
    // Each stream is kept a few items ahead of the device (events), so
    // neither waits for the host.
    hipStream_t st_pm;
    RR_HIP(rr::stream_create(&st_pm));
    SG guard_pm{st_pm};
    TickRunner T;
    if ((s = T.init(st, cus)) != RR_OK) return s;
    PmScheduler S;
    {
        uint32_t max_logn = 2;
        for (uint32_t i = 0; i < count; i++) {
            const uint64_t d = order && order[i] ? order[i] : seq_len_of(i) - 1;
            uint32_t l = 1;
            while (l < 63 && (1ull << l) < d + 1) l++;
            max_logn = std::max(max_logn, l);
        }
        if ((s = S.init(p, st_pm, max_logn, 128, 64)) != RR_OK) return s;
    }
    EventRing kq, pq;
    if ((s = kq.init(2)) != RR_OK || (s = pq.init(BATCH_PM_AHEAD)) != RR_OK) return s;
    hipEvent_t handoff;
    RR_HIP(hipEventCreateWithFlags(&handoff, hipEventDisableTiming));
    struct EG {
        hipEvent_t e;
        ~EG() { (void)hipEventDestroy(e); }
    } guard_ev{handoff};
    std::vector<Request *> active;
    std::vector<PmJob *> jobs;
    bool krylov_done = false, pm_idle = false;
    auto t2 = t1;
    for (;;) {
        bool issued = false;
        // requests whose set-up is done join the tick loop
        if (njoined < count) {
            for (uint32_t i = 0; i < count; i++) {
                if (joined[i]) continue;
                const int rd = su.ready[i].load(std::memory_order_acquire);
                if (rd == 2) {
                    set_error("matrix %u: %s", i, su.msg[i].c_str());
                    return su.st[i];
                }
                if (rd == 1) {
                    RR_HIP(hipStreamWaitEvent(st, su.ev[i], 0));
                    joined[i] = 1;
                    njoined++;
                    Request *r = reqs[i].get();
                    if (r->d >= PM_STREAM_D) {  // its PM-basis starts on the leading coefficients
                        if ((s = r->alloc_pm()) != RR_OK || (s = S.add(&r->pm)) != RR_OK) return s;
                        jobs.push_back(&r->pm);
                    }
                }
            }
            if (njoined == count) t1 = std::chrono::steady_clock::now();
        }
        std::vector<Request *> fresh;
        active.clear();
        uint64_t left = ~0ull;
        size_t words = 0;  // partial words of one tick
        for (uint32_t i = 0; i < count; i++) {
            if (!joined[i]) continue;  // reqs[i] may still be written by its set-up thread
            Request *r = reqs[i].get();
            if (r->stage != STAGE_KRYLOV || r->t >= r->L) continue;
            if (r->t == 0) {
                fresh.push_back(r);
                continue;
            }
            active.push_back(r);
            left = std::min<uint64_t>(left, r->L - r->t);
            words += (size_t)r->nwg * r->P;
        }
        if ((!fresh.empty() || !active.empty()) && kq.room()) {
            // requests at t = 0 take their projection tick alone; the others up to
            // BATCH_TICKS ticks per chunk, none past a sequence's end, partials within 64 M words
            if (!fresh.empty() && (s = T.tick(fresh, p, 1)) != RR_OK) return s;
            if (!active.empty()) {
                uint64_t n = std::min<uint64_t>(BATCH_TICKS, left);
                n = std::max<uint64_t>(1, std::min<uint64_t>(n, (64ull << 20) / std::max<size_t>(words, 1)));
                if ((s = T.tick(active, p, (uint32_t)n)) != RR_OK) return s;
            }
            if ((s = kq.push(st)) != RR_OK) return s;
            issued = true;
        }
        // complete sequences: G is built on st, the PM stream waits for it;
        // PM-basis jobs already running get the coefficients their next
        // operation waits for as soon as the sequence has them
        const size_t njobs = jobs.size();
        bool grown = false;
        for (uint32_t i = 0; i < count; i++) {
            Request *r = joined[i] ? reqs[i].get() : nullptr;
            if (!r || r->stage != STAGE_KRYLOV) continue;
            const bool running = r->G.ptr != nullptr;
            if (r->t >= r->L) {
                if ((s = r->start_pm(st, false)) != RR_OK) return s;
                if (running) grown = true;
                else jobs.push_back(&r->pm);
            } else if (running && r->pm.has_op && r->pm.op_need > r->pm.avail && r->t >= r->pm.op_need) {
                if ((s = r->grow_G(st)) != RR_OK) return s;
                grown = true;
            }
        }
        const bool handed = jobs.size() > njobs;
        if (handed || grown) {
            RR_HIP(hipEventRecord(handoff, st));
            RR_HIP(hipStreamWaitEvent(st_pm, handoff, 0));
            for (size_t j = njobs; j < jobs.size(); j++)
                if ((s = S.add(jobs[j])) != RR_OK) return s;
            pm_idle = false;
            issued = true;
        }
        if (!pm_idle && pq.room()) {
            bool launched = false;
            if ((s = S.round(jobs, &launched)) != RR_OK) return s;
            for (auto *J : jobs)
                if (J->err != RR_OK) return J->err;
            if (launched) {
                if ((s = pq.push(st_pm)) != RR_OK) return s;
                issued = true;
            } else {
                pm_idle = true;  // every joined job is done or waits for its sequence
            }
        }
        if (!krylov_done && njoined == count && active.empty() && fresh.empty() && !handed && !grown) {
            RR_HIP(hipStreamSynchronize(st));
            krylov_done = true;
            t2 = std::chrono::steady_clock::now();
        }
        if (krylov_done && pm_idle) break;
        if (!issued && !kq.retire() && !pq.retire()) std::this_thread::yield();  // streams ahead, or set-up pending
    }
    RR_HIP(hipStreamSynchronize(st_pm));
    uint64_t rounds = S.rounds, launches = S.launches;
    for (uint32_t i = 0; i < count; i++) {
        if ((s = reqs[i]->finish_rank(st_pm)) != RR_OK) return s;
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
// Autotuner: block width b (and with it rows per wavefront, 64/b) per matrix,
// for the batched and served modes.  There many matrices share the device,
// so what a width costs is the device time its work takes among the others
// (throughput), not its latency alone.  Measured once per process on the
// device, with copies of each probe running together:
//   c_b, f_b  Krylov time per tick and matrix, c_b per nonzero plus f_b:
//         TUNE_KCOPIES copies of a probe matrix with +-1/+-2 entries ticking
//         together, 16 ticks timed with events, at two sizes (TUNE_M rows);
//   pm_b(d) PM-basis time per job at orders 255 and 1023 for width b:
//         TUNE_COPIES jobs of random sequences in one batched scheduler run,
//         divided by the jobs, fitted as alpha_b d + beta_b d log2 d (leaves
//         are linear in d, the products of the d&c add d log d).
// Predicted device time of matrix i with width b:
//   L(b) * (nnz_i * c_b + f_b) + pm_b(d(b)),
//   L(b) = 2 min(m, n) / b + 8, d(b) = L(b) - 1.
constexpr uint32_t TUNE_KCOPIES = 32, TUNE_COPIES = 8;

namespace rr {

struct TuneProbe {
    bool done = false;
    double c[3] = {0, 0, 0}, pm[3] = {0, 0, 0};
    double pm_at[3][2] = {}, alpha[3] = {0, 0, 0}, beta[3] = {0, 0, 0}, f[3] = {0, 0, 0};
    double slot[3][2] = {};  // ms per tick per padded slot (nonzeros plus padding) at su = 3, 4
};
static const uint64_t TUNE_D[2] = {255, 1023};
static const uint64_t TUNE_M[2] = {24000, 3000};
constexpr double TUNE_MARGIN = 0.75;
static TuneProbe g_probe;
static const uint32_t TUNE_B[3] = {8, 16, 32};

static rr_status run_probe(uint32_t p, int cus, hipStream_t st) {
    uint64_t x = 0x9E3779B97F4A7C15ull;
    auto rnd = [&]() {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        return x;
    };
    hipEvent_t e0, e1;
    RR_HIP(hipEventCreate(&e0));
    RR_HIP(hipEventCreate(&e1));
    // ms per tick and copy: TUNE_KCOPIES copies of an m x 0.9m matrix with 11
    // +-1/+-2 entries per row ticking together
    auto krylov_probe = [&](uint64_t m, uint32_t b, uint32_t su, double *ms_tick, double *slots) -> rr_status {
        const uint64_t n = m * 9 / 10, k = 11;
        std::vector<uint64_t> rp(m + 1);
        std::vector<uint32_t> col(m * k), val(m * k);
        for (uint64_t r = 0; r < m; r++) {
            rp[r] = r * k;
            for (uint64_t e = 0; e < k; e++) {
                col[r * k + e] = (uint32_t)(rnd() % n);
                uint32_t mag = (rnd() & 3) == 0 ? 2 : 1;
                val[r * k + e] = (rnd() & 1) ? p - mag : mag;
            }
        }
        rp[m] = m * k;
        {  // padded slots of the two operands at this block size
            std::vector<uint32_t> r32(m + 1), ct(n + 1, 0);
            for (uint64_t r = 0; r <= m; r++) r32[r] = (uint32_t)rp[r];
            for (uint64_t e = 0; e < m * k; e++) ct[col[e] + 1]++;
            for (uint64_t c = 0; c < n; c++) ct[c + 1] += ct[c];
            *slots = (double)padded_slots(r32, ct, su);
        }
        std::vector<uint32_t> dr(m, 3), dc(n, 5), V(n * b);
        for (auto &v : V) v = (uint32_t)(rnd() % p);
        rr_csr A{m, n, rp.data(), col.data(), val.data()};
        rr_status s;
        std::vector<std::unique_ptr<Request>> R(TUNE_KCOPIES);
        std::vector<Request *> all;
        for (auto &r : R) {
            r.reset(new Request());
            if ((s = r->init(&A, p, b, dr.data(), dc.data(), V.data(), 40, 39, 2048, st, su)) != RR_OK) return s;
            all.push_back(r.get());
        }
        TickRunner T;
        if ((s = T.init(st, cus)) != RR_OK) return s;
        if ((s = T.tick(all, p)) != RR_OK) return s;  // t = 0 alone
        if ((s = T.tick(all, p, 3)) != RR_OK) return s;
        RR_HIP(hipEventRecord(e0, st));
        if ((s = T.tick(all, p, 16)) != RR_OK) return s;
        RR_HIP(hipEventRecord(e1, st));
        RR_HIP(hipEventSynchronize(e1));
        float ms = 0;
        RR_HIP(hipEventElapsedTime(&ms, e0, e1));
        *ms_tick = ms / 16.0 / TUNE_KCOPIES;
        return RR_OK;
    };
    rr_status s;
    for (int i = 0; i < 3; i++) {
        uint32_t b = TUNE_B[i];
        // cost per tick = c nnz + f: c from the difference of two sizes, f the rest
        double t_big = 0, t_small = 0, t_big4 = 0, sl_big = 0, sl_small = 0, sl_big4 = 0;
        if ((s = krylov_probe(TUNE_M[0], b, 3, &t_big, &sl_big)) != RR_OK) return s;
        if ((s = krylov_probe(TUNE_M[1], b, 3, &t_small, &sl_small)) != RR_OK) return s;
        if ((s = krylov_probe(TUNE_M[0], b, 4, &t_big4, &sl_big4)) != RR_OK) return s;
        const double nz_big = 11.0 * TUNE_M[0], nz_small = 11.0 * TUNE_M[1];
        g_probe.c[i] = std::max(0.0, (t_big - t_small) / (nz_big - nz_small));
        g_probe.f[i] = std::max(0.0, t_small - g_probe.c[i] * nz_small);
        // the same matrix at both block sizes: the tick time beyond f per padded slot
        g_probe.slot[i][0] = std::max(0.0, t_big - g_probe.f[i]) / sl_big;
        g_probe.slot[i][1] = std::max(0.0, t_big4 - g_probe.f[i]) / sl_big4;
        // PM-basis probe: TUNE_COPIES jobs of random sequences at each order of TUNE_D
        const uint64_t P = b * (b + 1) / 2;
        for (int o = 0; o < 2; o++) {
            const uint64_t d = TUNE_D[o];
            std::vector<uint32_t> seq((d + 1) * P * TUNE_COPIES);
            for (auto &v : seq) v = (uint32_t)(rnd() % p);
            DevBuf dseq;
            std::vector<DevBuf> G(TUNE_COPIES), M(TUNE_COPIES), dl(TUNE_COPIES);
            std::vector<PmJob> J(TUNE_COPIES);
            std::vector<PmJob *> jobs;
            if ((s = dseq.alloc(seq.size() * 4)) != RR_OK) return s;
            RR_HIP(hipMemcpyAsync(dseq.ptr, seq.data(), seq.size() * 4, hipMemcpyHostToDevice, st));
            for (uint32_t c = 0; c < TUNE_COPIES; c++) {
                if ((s = G[c].alloc((size_t)2 * b * b * (d + 1) * 4)) != RR_OK) return s;
                if ((s = M[c].alloc((size_t)4 * b * b * (d + 1) * 4)) != RR_OK) return s;
                if ((s = dl[c].alloc((size_t)2 * b * 8)) != RR_OK) return s;
                if ((s = build_G_device(dseq.as<uint32_t>() + (size_t)c * (d + 1) * P, b, d + 1,
                                        G[c].as<uint32_t>(), st)) != RR_OK)
                    return s;
                J[c].b = b;
                J[c].d = d;
                J[c].G = G[c].as<uint32_t>();
                J[c].M = M[c].as<uint32_t>();
                J[c].delta = dl[c].as<int64_t>();
                jobs.push_back(&J[c]);
            }
            RR_HIP(hipStreamSynchronize(st));
            auto t0 = std::chrono::steady_clock::now();
            if ((s = pm_basis_batch(p, jobs, st, nullptr, nullptr)) != RR_OK) return s;
            RR_HIP(hipStreamSynchronize(st));
            g_probe.pm_at[i][o] =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / TUNE_COPIES;
        }
        // pm(d) = alpha d + beta d log2 d through the two probes
        const double l0 = std::log2((double)TUNE_D[0]), l1 = std::log2((double)TUNE_D[1]);
        const double r0 = g_probe.pm_at[i][0] / TUNE_D[0], r1 = g_probe.pm_at[i][1] / TUNE_D[1];
        g_probe.beta[i] = std::max(0.0, (r1 - r0) / (l1 - l0));
        g_probe.alpha[i] = std::max(0.0, r0 - g_probe.beta[i] * l0);
        g_probe.pm[i] = g_probe.pm_at[i][0];
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    g_probe.done = true;
    return RR_OK;
}

}  // namespace rr

extern "C" {

rr_status rr_autotune_b(uint32_t count, const rr_csr *mats, uint32_t p, uint32_t *b_out, double *pred_ms) {
    return rr::guarded([&]() -> rr_status {
        clear_error();
        RR_CHECK(count == 0 || (mats && b_out), RR_EINVAL, "null argument");
        RR_CHECK(p == NTT_PRIME, RR_EINVAL, "the autotuner's PM-basis probe needs p = 2013265921");
        if (!g_probe.done) {
            int dev = 0, cus = 0;
            RR_HIP(hipGetDevice(&dev));
            RR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            hipStream_t st;
            RR_HIP(rr::stream_create(&st));
            rr_status s = run_probe(p, cus, st);
            rr::stream_destroy(st);
            if (s != RR_OK) return s;
        }
        // b = 16 unless another width is predicted TUNE_MARGIN faster: the probes
        // vary by 10-20 % between runs, and a width chosen on noise costs the
        // whole batch (b = 8 doubles a sequence's ticks)
        for (uint32_t i = 0; i < count; i++) {
            const rr_csr &A = mats[i];
            uint64_t nnz = A.row_ptr[A.m], mn = std::min(A.m, A.n);
            double t[3];
            for (int k = 0; k < 3; k++) {
                const double L = 2.0 * (double)mn / TUNE_B[k] + 8, d = L - 1;
                t[k] = d + 1 > (double)(1u << NTT_MAX_LOGN)
                           ? 1e300
                           : L * ((double)nnz * g_probe.c[k] + g_probe.f[k]) +
                                 d * (g_probe.alpha[k] + g_probe.beta[k] * std::log2(d));
            }
            int kb = 1;
            for (int k = 0; k < 3; k++)
                if (k != 1 && t[k] < TUNE_MARGIN * t[1] && t[k] < t[kb]) kb = k;
            b_out[i] = TUNE_B[kb];
            if (pred_ms) pred_ms[i] = t[kb];
        }
        return RR_OK;
    });
}

rr_status rr_autotune(uint32_t count, const rr_csr *mats, uint32_t p, rr_tuning *out) {
    return rr::guarded([&]() -> rr_status {
        clear_error();
        RR_CHECK(count == 0 || (mats && out), RR_EINVAL, "null argument");
        std::vector<uint32_t> b(count);
        std::vector<double> pred(count);
        rr_status s = rr_autotune_b(count, mats, p, b.data(), pred.data());
        if (s != RR_OK) return s;
        for (uint32_t i = 0; i < count; i++) {
            const rr_csr &A = mats[i];
            RR_CHECK(A.row_ptr && A.row_ptr[0] == 0 && A.row_ptr[A.m] < (1ull << 32), RR_EINVAL,
                     "matrix %u: bad row_ptr", i);
            const uint64_t nnz = A.row_ptr[A.m];
            std::vector<uint32_t> rp(A.m + 1), ct(A.n + 1, 0);
            for (uint64_t r = 0; r <= A.m; r++) rp[r] = (uint32_t)A.row_ptr[r];
            for (uint64_t e = 0; e < nnz; e++) {
                RR_CHECK(A.col[e] < A.n, RR_EINVAL, "matrix %u: column index out of range", i);
                ct[A.col[e] + 1]++;
            }
            for (uint64_t c = 0; c < A.n; c++) ct[c + 1] += ct[c];
            const int k = b[i] == 8 ? 0 : b[i] == 16 ? 1 : 2;
            const double L = 2.0 * (double)std::min(A.m, A.n) / b[i] + 8;
            double t[2];
            for (int u = 0; u < 2; u++) t[u] = L * g_probe.slot[k][u] * (double)padded_slots(rp, ct, 3 + u);
            // SU = 4 unless SU = 3 is predicted at least 15% faster: the probes'
            // per-slot costs vary 10-20% between runs, and choices of SU = 3
            // inside that spread made a batch slower
            const int uk = t[0] < 0.85 * t[1] ? 0 : 1;
            out[i].b = b[i];
            out[i].su = 3 + uk;
            // the b model's Krylov term replaced by the slot model's at the chosen block size
            const double kr = L * ((double)nnz * g_probe.c[k] + g_probe.f[k]);
            out[i].pred_ms = pred[i] - kr + L * g_probe.f[k] + t[uk];
        }
        return RR_OK;
    });
}

rr_status rr_autotune_probe(double *c_ms_per_nnz, double *pm_ms_d255) {
    return rr::guarded([&]() -> rr_status {
        RR_CHECK(c_ms_per_nnz && pm_ms_d255, RR_EINVAL, "null argument");
        RR_CHECK(g_probe.done, RR_ESTATE, "autotuner has not probed yet");
        for (int k = 0; k < 3; k++) {
            c_ms_per_nnz[k] = g_probe.c[k];
            pm_ms_d255[k] = g_probe.pm[k];
        }
        return RR_OK;
    });
}

}  // extern "C"
