// setup.hip -- device-side set-up of a matrix's sliced operands.
//
// Replaces the host passes of O(nnz) that prepare A for the Krylov passes:
// validation of the entries (src/matrices.rs:842-870 guards the reference's
// types), the transpose (src/matrices.rs:734-778, row order kept within each
// column) and the copy of every nonzero into its wavefront's column-block
// stream.  The host keeps the O(rows) planning: row lengths (from row_ptr and
// the device's column histogram), the length-sorted slices and their
// longest-first assignment to wavefronts (build_slices).  The preconditioners
// are applied to vectors, never to A, so A is uploaded once as given.
//
// The device path serves matrices with no row of A or A^T long enough to be
// cut into segments (uniform, batched and served workloads); a matrix with
// such rows (a Zipf head) is laid out on the host, where the segments are cut
// at the XCD eighths of the gathered vector.  Both paths give the same
// layout for the same matrix, and every result is exact either way.
#include "common.hpp"
#include "spmm.hpp"
#include <algorithm>
#include <vector>

namespace rr {

// Entry checks and the column histogram: the first entry with a column >= n,
// with a value >= p, and with a value that is not a signed residue in
// [-8, 7] (flags[0..2], ~0 when none); cnt[c] = entries of column c.
__global__ void k_check_csr(const uint32_t *__restrict__ col, const uint32_t *__restrict__ val, uint64_t nnz,
                            uint32_t n, uint32_t p, uint32_t *cnt, unsigned long long *flags) {
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t c = col[e], v = val[e];
        if (c >= n) atomicMin(flags + 0, (unsigned long long)e);
        else atomicAdd(cnt + c, 1u);
        if (v >= p) atomicMin(flags + 1, (unsigned long long)e);
        if (!(v <= 7 || v >= p - 8)) atomicMin(flags + 2, (unsigned long long)e);
    }
}

__device__ __forceinline__ uint32_t value_code(uint32_t v, uint32_t p) {
    const int32_t sv = v <= 7 ? (int32_t)v : (int32_t)v - (int32_t)p;
    return ((uint32_t)sv & 0xFu) << 28;
}

// Packed column words of A: column | 4-bit signed value << 28.
__global__ void k_pack_cols(const uint32_t *__restrict__ col, const uint32_t *__restrict__ val, uint64_t nnz,
                            uint32_t p, uint32_t *__restrict__ out) {
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (uint64_t)gridDim.x * blockDim.x)
        out[e] = col[e] | value_code(val[e], p);
}

// A^T's entries into their columns' ranges [rpt[c], rpt[c + 1]) (one thread
// per row of A; fill: zeroed per-column counters).  The order within a
// column is restored to row order by k_sort_cols.
__global__ void k_transpose_scatter(const uint32_t *__restrict__ rp, const uint32_t *__restrict__ col,
                                    const uint32_t *__restrict__ val, uint32_t m, bool packed, uint32_t p,
                                    const uint32_t *__restrict__ rpt, uint32_t *fill, uint32_t *__restrict__ colt,
                                    uint32_t *__restrict__ valt) {
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < m; r += gridDim.x * blockDim.x) {
        for (uint32_t e = rp[r]; e < rp[r + 1]; e++) {
            const uint32_t c = col[e], v = val[e];
            const uint32_t pos = rpt[c] + atomicAdd(fill + c, 1u);
            if (packed) {
                colt[pos] = r | value_code(v, p);
            } else {
                colt[pos] = r;
                valt[pos] = v;
            }
        }
    }
}

// Each column's entries in row order (insertion sort, one thread per column;
// the device path only runs when every column is short).
__global__ void k_sort_cols(const uint32_t *__restrict__ rpt, uint32_t n, bool packed, uint32_t *colt,
                            uint32_t *valt) {
    const uint32_t mask = packed ? 0x0FFFFFFFu : 0xFFFFFFFFu;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        const uint32_t b = rpt[c], e = rpt[c + 1];
        for (uint32_t i = b + 1; i < e; i++) {
            const uint32_t w = colt[i], v = packed ? 0u : valt[i], key = w & mask;
            uint32_t j = i;
            while (j > b && (colt[j - 1] & mask) > key) {
                colt[j] = colt[j - 1];
                if (!packed) valt[j] = valt[j - 1];
                j--;
            }
            colt[j] = w;
            if (!packed) valt[j] = v;
        }
    }
}

// Column blocks of row slices: slice s (sl[s] = {first row in the length
// order, first block}) holds rows shortr[first + g], g < RPW; nonzero e of
// group g goes to word (block + e / SU) * RPW * SU + g * SU + e % SU.  One
// thread per (slice, group); cw / vw are zero beforehand (padding).
__global__ void k_fill_slices(const uint32_t *__restrict__ rp, const uint32_t *__restrict__ colw,
                              const uint32_t *__restrict__ vals, const uint32_t *__restrict__ shortr, uint32_t nshort,
                              const uint2 *__restrict__ sl, uint32_t nsl, uint32_t RPW, uint32_t SU,
                              uint32_t *__restrict__ cw, uint32_t *__restrict__ vw) {
    const uint64_t units = (uint64_t)nsl * RPW;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < units; t += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t s = (uint32_t)(t / RPW), g = (uint32_t)(t % RPW);
        const uint2 d = sl[s];
        const uint32_t k = d.x + g;
        if (k >= nshort) continue;
        const uint32_t r = shortr[k], beg = rp[r], len = rp[r + 1] - beg;
        for (uint32_t e = 0; e < len; e++) {
            const size_t w = ((size_t)d.y + e / SU) * RPW * SU + g * SU + e % SU;
            cw[w] = colw[beg + e];
            if (vals) vw[w] = vals[beg + e];
        }
    }
}

static unsigned grid_for(uint64_t items) { return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((items + 255) / 256, 65536)); }

void launch_fill_slices(const DevOperand &op, const uint32_t *shortr, uint32_t nshort, const uint2 *sl, uint32_t nsl,
                        uint32_t RPW, uint32_t SU, uint32_t *cw, uint32_t *vw, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_slices, dim3(grid_for((uint64_t)nsl * RPW)), dim3(256), 0, st, op.rp, op.colw, op.vals,
                       shortr, nshort, sl, nsl, RPW, SU, cw, vw);
}

// The first index i with v[i] >= p (atomic minimum; a thread stops at its
// first such index, all later ones of its stride being larger).
__global__ void k_first_not_below(const uint32_t *__restrict__ v, uint64_t n, uint32_t p,
                                  unsigned long long *first) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        if (v[i] >= p) {
            atomicMin(first, (unsigned long long)i);
            break;
        }
}

// Rows of b words to rows of B words, zero past column b.
__global__ void k_pad_rows(const uint32_t *__restrict__ in, uint32_t *__restrict__ out, uint64_t nrows, uint32_t b,
                           uint32_t B) {
    const uint64_t n = nrows * B;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t r = i / B;
        const uint32_t j = (uint32_t)(i - r * B);
        out[i] = j < b ? in[r * b + j] : 0u;
    }
}

rr_status upload_block_vector(DevBuf &out, const uint32_t *V, uint64_t n, uint32_t b, uint32_t B, uint32_t p,
                              const char *name, hipStream_t st) {
    RR_CHECK(b >= 1 && b <= B, RR_EINVAL, "block width b=%u above the row stride %u", b, B);
    rr_status s;
    if ((s = out.alloc((size_t)n * B * 4)) != RR_OK) return s;
    DevBuf raw, flag;
    uint32_t *src = out.as<uint32_t>();
    if (b != B) {
        if ((s = raw.alloc((size_t)n * b * 4)) != RR_OK) return s;
        src = raw.as<uint32_t>();
    }
    if ((s = flag.alloc(sizeof(unsigned long long))) != RR_OK) return s;
    RR_HIP(hipMemcpyAsync(src, V, (size_t)n * b * 4, hipMemcpyHostToDevice, st));
    RR_HIP(hipMemsetAsync(flag.ptr, 0xFF, sizeof(unsigned long long), st));
    hipLaunchKernelGGL(k_first_not_below, dim3(grid_for(n * b)), dim3(256), 0, st, src, n * b, p,
                       flag.as<unsigned long long>());
    if (b != B)
        hipLaunchKernelGGL(k_pad_rows, dim3(grid_for(n * B)), dim3(256), 0, st, src, out.as<uint32_t>(), n, b, B);
    RR_HIP(hipGetLastError());
    unsigned long long first = 0;
    RR_HIP(hipMemcpyAsync(&first, flag.ptr, sizeof first, hipMemcpyDeviceToHost, st));
    RR_HIP(hipStreamSynchronize(st));  // also before raw goes back to the cache
    RR_CHECK(first == ~0ull, RR_EINVAL, "%s[%llu] >= p", name, first);
    return RR_OK;
}

template <typename T>
static rr_status upload_async(DevBuf &buf, const T *host, size_t count, hipStream_t st) {
    rr_status s = buf.alloc(count * sizeof(T));
    if (s != RR_OK) return s;
    if (count) RR_HIP(hipMemcpyAsync(buf.ptr, host, count * sizeof(T), hipMemcpyHostToDevice, st));
    return RR_OK;
}

rr_status setup_operands(const rr_csr *A, uint32_t p, uint32_t B, uint32_t su_req, const NwOf &nw_of,
                         const uint32_t *row_scale, const uint32_t *col_scale2, SliceMat &sa, SliceMat &sat,
                         OperandInfo *info, hipStream_t st) {
    const uint64_t m = A->m, n = A->n;
    RR_CHECK(A->row_ptr && A->row_ptr[0] == 0, RR_EINVAL, "row_ptr[0] must be 0");
    for (uint64_t r = 0; r < m; r++)
        RR_CHECK(A->row_ptr[r + 1] >= A->row_ptr[r], RR_EINVAL, "row_ptr not monotone at row %llu",
                 (unsigned long long)r);
    const uint64_t nnz = A->row_ptr[m];
    RR_CHECK(nnz < (1ull << 32), RR_EINVAL, "nnz=%llu must be < 2^32", (unsigned long long)nnz);
    RR_CHECK(nnz == 0 || (A->col && A->val), RR_EINVAL, "null col/val");

    // entries to the device, checked there with the column histogram
    std::vector<uint32_t> rp(m + 1);
    for (uint64_t r = 0; r <= m; r++) rp[r] = (uint32_t)A->row_ptr[r];
    DevBuf d_rp, d_col, d_val, d_cnt, d_flags;
    rr_status s;
    if ((s = upload_async(d_rp, rp.data(), rp.size(), st)) != RR_OK) return s;
    if ((s = upload_async(d_col, A->col, nnz, st)) != RR_OK) return s;
    if ((s = upload_async(d_val, A->val, nnz, st)) != RR_OK) return s;
    if ((s = d_cnt.alloc((size_t)n * 4)) != RR_OK) return s;
    if ((s = d_flags.alloc(3 * sizeof(unsigned long long))) != RR_OK) return s;
    RR_HIP(hipMemsetAsync(d_cnt.ptr, 0, (size_t)n * 4, st));
    RR_HIP(hipMemsetAsync(d_flags.ptr, 0xFF, 3 * sizeof(unsigned long long), st));
    if (nnz)
        hipLaunchKernelGGL(k_check_csr, dim3(grid_for(nnz)), dim3(256), 0, st, d_col.as<uint32_t>(),
                           d_val.as<uint32_t>(), nnz, (uint32_t)n, p, d_cnt.as<uint32_t>(),
                           d_flags.as<unsigned long long>());
    RR_HIP(hipGetLastError());
    std::vector<uint32_t> cnt(n);
    unsigned long long flags[3];
    RR_HIP(hipMemcpyAsync(cnt.data(), d_cnt.ptr, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    RR_HIP(hipMemcpyAsync(flags, d_flags.ptr, sizeof flags, hipMemcpyDeviceToHost, st));
    RR_HIP(hipStreamSynchronize(st));
    RR_CHECK(flags[0] == ~0ull, RR_EINVAL, "col[%llu]=%u >= n", flags[0], A->col[flags[0]]);
    RR_CHECK(flags[1] == ~0ull, RR_EINVAL, "val[%llu]=%u >= p", flags[1], A->val[flags[1]]);
    const bool packed = flags[2] == ~0ull && m < (1ull << 28) && n < (1ull << 28);

    std::vector<uint32_t> rpt(n + 1, 0);
    for (uint64_t c = 0; c < n; c++) rpt[c + 1] = rpt[c] + cnt[c];
    const uint32_t su = su_req ? su_req : choose_su(rp, rpt);
    const uint32_t nw = nw_of(packed, su);
    uint32_t max_row = 0, max_col = 0;
    for (uint64_t r = 0; r < m; r++) max_row = std::max(max_row, rp[r + 1] - rp[r]);
    for (uint64_t c = 0; c < n; c++) max_col = std::max(max_col, cnt[c]);
    info->packed = packed;
    info->su = su;
    info->nw = nw;
    info->on_device = max_row <= long_row_threshold(nnz, m, B, su, nw) && max_col <= long_row_threshold(nnz, n, B, su, nw);

    if (!info->on_device) {  // rows to cut into segments: the host lays out both operands
        std::vector<uint32_t> hrp, colw, vals, hrpt, colt, valt;
        build_operands(A, p, packed, hrp, colw, vals, hrpt, colt, valt);
        if ((s = build_slices(sa, m, hrp, colw, vals, packed, B, su, nw, row_scale, n)) != RR_OK) return s;
        return build_slices(sat, n, hrpt, colt, valt, packed, B, su, nw, col_scale2, m);
    }

    // A's column words, then A^T by scatter and per-column row order
    DevBuf d_colw, d_rpt, d_fill, d_colt, d_valt;
    const uint32_t *colw = d_col.as<uint32_t>();
    if (packed) {
        if ((s = d_colw.alloc((size_t)nnz * 4)) != RR_OK) return s;
        if (nnz)
            hipLaunchKernelGGL(k_pack_cols, dim3(grid_for(nnz)), dim3(256), 0, st, d_col.as<uint32_t>(),
                               d_val.as<uint32_t>(), nnz, p, d_colw.as<uint32_t>());
        colw = d_colw.as<uint32_t>();
    }
    if ((s = upload_async(d_rpt, rpt.data(), rpt.size(), st)) != RR_OK) return s;
    if ((s = d_fill.alloc((size_t)n * 4)) != RR_OK) return s;
    if ((s = d_colt.alloc((size_t)nnz * 4)) != RR_OK) return s;
    if ((s = d_valt.alloc(packed ? 4 : (size_t)nnz * 4)) != RR_OK) return s;
    RR_HIP(hipMemsetAsync(d_fill.ptr, 0, (size_t)n * 4, st));
    hipLaunchKernelGGL(k_transpose_scatter, dim3(grid_for(m)), dim3(256), 0, st, d_rp.as<uint32_t>(),
                       d_col.as<uint32_t>(), d_val.as<uint32_t>(), (uint32_t)m, packed, p, d_rpt.as<uint32_t>(),
                       d_fill.as<uint32_t>(), d_colt.as<uint32_t>(), d_valt.as<uint32_t>());
    hipLaunchKernelGGL(k_sort_cols, dim3(grid_for(n)), dim3(256), 0, st, d_rpt.as<uint32_t>(), (uint32_t)n, packed,
                       d_colt.as<uint32_t>(), d_valt.as<uint32_t>());
    RR_HIP(hipGetLastError());
    const DevOperand opa{d_rp.as<uint32_t>(), colw, packed ? nullptr : d_val.as<uint32_t>()};
    const DevOperand opt{d_rpt.as<uint32_t>(), d_colt.as<uint32_t>(), packed ? nullptr : d_valt.as<uint32_t>()};
    const std::vector<uint32_t> none;
    if ((s = build_slices(sa, m, rp, none, none, packed, B, su, nw, row_scale, n, &opa, st)) != RR_OK) return s;
    if ((s = build_slices(sat, n, rpt, none, none, packed, B, su, nw, col_scale2, m, &opt, st)) != RR_OK) return s;
    // the temporaries go back to the cache, reused only once this stream's work is done
    RR_HIP(hipStreamSynchronize(st));
    return RR_OK;
}

}  // namespace rr
