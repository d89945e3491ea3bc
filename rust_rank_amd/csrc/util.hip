// util.hip -- the device memory cache behind DevBuf, and measurement
// helpers: a plain device copy kernel of known size, used to calibrate
// hardware byte counters (FETCH_SIZE / WRITE_SIZE) and to measure achievable
// HBM bandwidth on this GPU.
#include "common.hpp"
#include <deque>
#include <map>
#include <mutex>
#include <vector>

namespace rr {

namespace {
struct FreeBlock {
    void *ptr;
    uint64_t gen;  // free generation: safe to reuse once a marker recorded after it has completed
};
struct DevCache {
    std::map<size_t, std::deque<FreeBlock>> free;  // per size class, oldest first
    uint64_t gen_freed = 0, gen_synced = 0;
    // markers in flight: one event per library stream of this device, recorded
    // after every free up to gen_marker
    std::vector<hipEvent_t> marker;
    size_t nmark = 0;  // events of the live marker (one per stream it was recorded on)
    uint64_t gen_marker = 0;
    bool marker_live = false;
};
std::mutex g_cache_mu;
std::map<int, DevCache> g_cache;
std::map<hipStream_t, int> g_streams;  // library streams and their devices (under g_cache_mu)

size_t size_class(size_t n) {
    if (n <= 256) return 256;
    int k = 63 - __builtin_clzll((unsigned long long)(n - 1));  // n in (2^k, 2^(k+1)]
    if (k < 20) return (size_t)1 << (k + 1);
    const size_t step = (size_t)1 << (k - 3);
    return (n + step - 1) / step * step;
}

// Advances gen_synced when the markers in flight have completed, and puts
// new markers on the device's streams when frees are not covered yet
// (caller holds g_cache_mu; the device is current).  Never waits.
void advance_markers_(DevCache &D, int dev) {
    if (D.marker_live) {
        for (size_t i = 0; i < D.nmark; i++)
            if (hipEventQuery(D.marker[i]) != hipSuccess) return;  // still running (or an error: try again later)
        D.gen_synced = std::max(D.gen_synced, D.gen_marker);
        D.marker_live = false;
    }
    if (D.gen_freed <= D.gen_synced) return;
    // one event per library stream of the device (streams destroyed since the
    // frees waited for their work in stream_destroy)
    size_t k = 0;
    for (auto &kv : g_streams) {
        if (kv.second != dev) continue;
        if (k == D.marker.size()) {
            hipEvent_t e;
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return;
            D.marker.push_back(e);
        }
        if (hipEventRecord(D.marker[k], kv.first) != hipSuccess) return;
        k++;
    }
    D.nmark = k;
    D.gen_marker = D.gen_freed;
    D.marker_live = true;
}

void advance_markers(DevCache &D, int dev) {
    // leaves the thread's last-error state as it found it when it was clean
    const bool clean = hipPeekAtLastError() == hipSuccess;
    advance_markers_(D, dev);
    if (clean) (void)hipGetLastError();
}
}  // namespace

hipError_t stream_create(hipStream_t *st) {
    hipError_t e = hipStreamCreateWithFlags(st, hipStreamNonBlocking);
    if (e != hipSuccess) return e;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> g(g_cache_mu);
    g_streams[*st] = dev;
    return hipSuccess;
}

void stream_destroy(hipStream_t st) {
    if (!st) return;
    (void)hipStreamSynchronize(st);  // nothing on it can use a freed block any more
    {
        std::lock_guard<std::mutex> g(g_cache_mu);
        g_streams.erase(st);
    }
    (void)hipStreamDestroy(st);
}

void *dev_alloc(size_t bytes, size_t *cls, int *device, hipError_t *err) {
    int dev = 0;
    if ((*err = hipGetDevice(&dev)) != hipSuccess) return nullptr;
    const size_t c = size_class(bytes);
    *cls = c;
    *device = dev;
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        DevCache &D = g_cache[dev];
        auto it = D.free.find(c);
        if (it != D.free.end() && !it->second.empty()) {
            if (it->second.front().gen > D.gen_synced) advance_markers(D, dev);
            if (it->second.front().gen <= D.gen_synced) {
                void *p = it->second.front().ptr;
                it->second.pop_front();
                return p;
            }
        }
    }
    void *p = nullptr;
    *err = hipMalloc(&p, c);
    if (*err == hipErrorOutOfMemory || *err == hipErrorMemoryAllocation) {
        (void)hipGetLastError();
        dev_cache_release(dev);
        *err = hipMalloc(&p, c);
    }
    if (*err != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return p;
}

void dev_free(void *ptr, size_t cls, int device) {
    std::lock_guard<std::mutex> g(g_cache_mu);
    DevCache &D = g_cache[device];
    D.free[cls].push_back(FreeBlock{ptr, ++D.gen_freed});
}

void dev_cache_release(int device) {
    std::vector<void *> blocks;
    {
        std::lock_guard<std::mutex> g(g_cache_mu);
        DevCache &D = g_cache[device];
        for (auto &kv : D.free)
            for (const FreeBlock &fb : kv.second) blocks.push_back(fb.ptr);
        D.free.clear();
        D.gen_synced = D.gen_freed;  // after the synchronisation below nothing freed is in use
    }
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(device);
    (void)hipDeviceSynchronize();
    for (void *p : blocks) (void)hipFree(p);
    (void)hipSetDevice(cur);
}

__global__ __launch_bounds__(256) void k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Gather probe: row indices uniform in [0, rows) from a SplitMix64 hash of the position.
__global__ __launch_bounds__(256) void k_gather_index(uint32_t *idx, uint64_t n, uint64_t rows, uint64_t seed) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        idx[i] = (uint32_t)(z % rows);
    }
}

// Each group of 16 lanes reads one 256-byte row (16 x 16 bytes) per index,
// four indices in flight per group; the XOR of what it read is stored so the
// loads stay live.  With val != nullptr (the Krylov pass's traffic mix on
// rows of 20 entries) every gather also reads a 4-byte value, and each group
// writes one 256-byte row of Y (nout rows) per 20 gathers: every fifth
// iteration of its loop.
__global__ __launch_bounds__(256) void k_gather_rows(const uint4 *__restrict__ X, const uint32_t *__restrict__ idx,
                                                     const uint32_t *__restrict__ val, uint64_t n, uint4 *__restrict__ Y,
                                                     uint64_t nout, uint32_t *out) {
    const uint32_t lane = threadIdx.x & 63, g = lane >> 4, q = lane & 15;
    const uint64_t wave = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6), nwave = (uint64_t)gridDim.x * 4;
    uint32_t acc = 0;
    uint4 o = make_uint4(0u, 0u, 0u, 0u);
    uint32_t k = 0;
    for (uint64_t base = wave * 16; base < n; base += nwave * 16, k++) {
        uint32_t r[4], w[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint64_t i = base + u * 4 + g;
            r[u] = i < n ? idx[i] : 0u;
            w[u] = (val && i < n) ? val[i] : 1u;
        }
        uint4 v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) v[u] = X[(uint64_t)r[u] * 16 + q];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            o.x += v[u].x * w[u];
            o.y += v[u].y * w[u];
            o.z += v[u].z * w[u];
            o.w += v[u].w * w[u];
        }
        const uint64_t row = base / 16 / 5 * 4 + g;  // < n / 20 + 4
        if (Y && k % 5 == 4 && row < nout) Y[row * 16 + q] = o;
    }
    acc = o.x ^ o.y ^ o.z ^ o.w;
    out[(uint64_t)blockIdx.x * 256 + threadIdx.x] = acc;
}

}  // namespace rr

using namespace rr;

extern "C" {

rr_status rr_gather_probe(uint64_t rows, uint64_t gathers, uint32_t iters, int pass_mix, double *ms_out,
                          double *gbs_out) {
    return rr::guarded([&]() -> rr_status {
        clear_error();
        RR_CHECK(rows >= 1 && rows <= (1ull << 32) && gathers >= 1 && iters >= 1 && ms_out && gbs_out, RR_EINVAL,
                 "bad gather probe args");
        int dev = 0, cus = 0;
        RR_HIP(hipGetDevice(&dev));
        RR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        const uint32_t grid = (uint32_t)cus * 16;
        const uint64_t nout = gathers / 20 + 64;  // one row per 20 gathers
        DevBuf X, idx, out, val, Y;
        rr_status s;
        if ((s = X.alloc(rows * 256)) != RR_OK) return s;
        if ((s = idx.alloc(gathers * 4)) != RR_OK) return s;
        if ((s = out.alloc((size_t)grid * 256 * 4)) != RR_OK) return s;
        if (pass_mix) {
            if ((s = val.alloc(gathers * 4)) != RR_OK) return s;
            if ((s = Y.alloc(nout * 256)) != RR_OK) return s;
        }
        hipStream_t st;
        RR_HIP(stream_create(&st));
        struct S {
            hipStream_t s;
            ~S() { stream_destroy(s); }
        } guard{st};
        RR_HIP(hipMemsetAsync(X.ptr, 1, rows * 256, st));
        if (pass_mix) RR_HIP(hipMemsetAsync(val.ptr, 3, gathers * 4, st));
        hipLaunchKernelGGL(k_gather_index, dim3(grid), dim3(256), 0, st, idx.as<uint32_t>(), gathers, rows, 12345ull);
        const uint32_t *vp = pass_mix ? val.as<uint32_t>() : nullptr;
        uint4 *yp = pass_mix ? Y.as<uint4>() : nullptr;
        hipLaunchKernelGGL(k_gather_rows, dim3(grid), dim3(256), 0, st, X.as<uint4>(), idx.as<uint32_t>(), vp,
                           gathers, yp, nout, out.as<uint32_t>());
        hipEvent_t e0, e1;
        RR_HIP(hipEventCreate(&e0));
        RR_HIP(hipEventCreate(&e1));
        RR_HIP(hipEventRecord(e0, st));
        for (uint32_t i = 0; i < iters; i++)
            hipLaunchKernelGGL(k_gather_rows, dim3(grid), dim3(256), 0, st, X.as<uint4>(), idx.as<uint32_t>(), vp,
                               gathers, yp, nout, out.as<uint32_t>());
        RR_HIP(hipEventRecord(e1, st));
        RR_HIP(hipEventSynchronize(e1));
        float ms = 0;
        RR_HIP(hipEventElapsedTime(&ms, e0, e1));
        *ms_out = ms / iters;
        const double bytes = (double)gathers * (pass_mix ? 264.0 + 256.0 / 20.0 : 260.0);
        *gbs_out = bytes / (*ms_out * 1e-3) / 1e9;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return RR_OK;
    });
}

rr_status rr_release_device_memory(void) {
    return rr::guarded([&]() -> rr_status {
        clear_error();
        int dev = 0;
        RR_HIP(hipGetDevice(&dev));
        dev_cache_release(dev);
        return RR_OK;
    });
}

// Copies `bytes` (multiple of 16) device-to-device with k_copy, `iters` times;
// ms_out receives the average kernel time and gbs_out read+write GB/s.
rr_status rr_copy_probe(uint64_t bytes, uint32_t iters, double *ms_out, double *gbs_out) {
    return rr::guarded([&]() -> rr_status {
        clear_error();
        RR_CHECK(bytes >= 16 && bytes % 16 == 0 && iters >= 1 && ms_out && gbs_out, RR_EINVAL, "bad copy probe args");
        DevBuf a, b;
        rr_status s;
        if ((s = a.alloc(bytes)) != RR_OK) return s;
        if ((s = b.alloc(bytes)) != RR_OK) return s;
        hipStream_t st;
        RR_HIP(stream_create(&st));
        RR_HIP(hipMemsetAsync(a.ptr, 1, bytes, st));
        int dev = 0, cus = 0;
        RR_HIP(hipGetDevice(&dev));
        RR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        size_t n = bytes / 16;
        dim3 grid((unsigned)std::min<size_t>((n + 255) / 256, (size_t)cus * 16));
        hipLaunchKernelGGL(k_copy, grid, dim3(256), 0, st, a.as<uint4>(), b.as<uint4>(), n);
        hipEvent_t e0, e1;
        RR_HIP(hipEventCreate(&e0));
        RR_HIP(hipEventCreate(&e1));
        RR_HIP(hipEventRecord(e0, st));
        for (uint32_t i = 0; i < iters; i++)
            hipLaunchKernelGGL(k_copy, grid, dim3(256), 0, st, a.as<uint4>(), b.as<uint4>(), n);
        RR_HIP(hipEventRecord(e1, st));
        RR_HIP(hipEventSynchronize(e1));
        float ms = 0;
        RR_HIP(hipEventElapsedTime(&ms, e0, e1));
        *ms_out = ms / iters;
        *gbs_out = 2.0 * (double)bytes / (*ms_out * 1e-3) / 1e9;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        stream_destroy(st);
        return RR_OK;
    });
}

}  // extern "C"
