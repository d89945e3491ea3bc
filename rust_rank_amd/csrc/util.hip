// util.hip -- measurement helpers: a plain device copy kernel of known size,
// used to calibrate hardware byte counters (FETCH_SIZE / WRITE_SIZE) and to
// measure achievable HBM bandwidth on this GPU.
#include "common.hpp"

namespace rr {

__global__ __launch_bounds__(256) void k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

}  // namespace rr

using namespace rr;

extern "C" {

// Copies `bytes` (multiple of 16) device-to-device with k_copy, `iters` times;
// ms_out receives the average kernel time and gbs_out read+write GB/s.
rr_status rr_copy_probe(uint64_t bytes, uint32_t iters, double *ms_out, double *gbs_out) {
    clear_error();
    RR_CHECK(bytes >= 16 && bytes % 16 == 0 && iters >= 1 && ms_out && gbs_out, RR_EINVAL, "bad copy probe args");
    DevBuf a, b;
    rr_status s;
    if ((s = a.alloc(bytes)) != RR_OK) return s;
    if ((s = b.alloc(bytes)) != RR_OK) return s;
    hipStream_t st;
    RR_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
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
    for (uint32_t i = 0; i < iters; i++) hipLaunchKernelGGL(k_copy, grid, dim3(256), 0, st, a.as<uint4>(), b.as<uint4>(), n);
    RR_HIP(hipEventRecord(e1, st));
    RR_HIP(hipEventSynchronize(e1));
    float ms = 0;
    RR_HIP(hipEventElapsedTime(&ms, e0, e1));
    *ms_out = ms / iters;
    *gbs_out = 2.0 * (double)bytes / (*ms_out * 1e-3) / 1e9;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    return RR_OK;
}

}  // extern "C"
