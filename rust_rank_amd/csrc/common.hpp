// common.hpp -- shared host/device helpers for the rank engine (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstdarg>
#include <string>
#include <new>
#include <stdexcept>
#include "../../include/rank_engine.h"

namespace rr {

void set_error(const char *fmt, ...);
void clear_error();

// The body of a C entry point: a host exception (an allocation that fails, a
// size beyond what a container can hold) becomes a status and a message for
// the calling thread instead of unwinding through the C ABI.
template <class F>
rr_status guarded(F &&body) {
    try {
        return body();
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed");
        return RR_ENOMEM;
    } catch (const std::length_error &e) {
        set_error("host buffer too large (%s)", e.what());
        return RR_ENOMEM;
    } catch (const std::exception &e) {
        set_error("host error: %s", e.what());
        return RR_EINVAL;
    }
}

// Loads and stores through a pointer marked global: the compiler emits
// global_* (in-order VMEM) instead of flat_* instructions, whose completion
// it cannot order against other memory operations and so waits out in full.
#define GLOBAL_AS __attribute__((address_space(1)))
typedef uint32_t gv2u __attribute__((ext_vector_type(2)));
typedef uint32_t gv4u __attribute__((ext_vector_type(4)));
template <typename T>
__device__ __forceinline__ const GLOBAL_AS T *as_global(const T *p) {
    return (const GLOBAL_AS T *)p;
}
template <typename T>
__device__ __forceinline__ GLOBAL_AS T *as_global(T *p) {
    return (GLOBAL_AS T *)p;
}

// ---------------------------------------------------------------------------
// Exact arithmetic mod p < 2^31.  Integer-valued doubles are exact below 2^53;
// FP64 FMA is full rate on CDNA4, 32-bit integer multiplies are not.

struct ModP {
    uint32_t p;
    double pd;   // (double)p
    double inv;  // 1.0 / p
    static ModP make(uint32_t p) { return ModP{p, (double)p, 1.0 / (double)p}; }
};

// a: integer-valued double with |a| < 2^53  ->  a mod p in [0, p)
__device__ __forceinline__ double fold(double a, double pd, double inv) {
    double q = floor(a * inv);
    double r = fma(-q, pd, a);  // exact: the true result is a small integer
    r = r < 0.0 ? r + pd : r;
    r = r >= pd ? r - pd : r;
    return r;
}

// a * b mod p for a, b < p < 2^31
__device__ __forceinline__ uint32_t mulmod(uint32_t a, uint32_t b, uint32_t p, double inv) {
    double x = (double)a * (double)b;  // relative error <= 2^-53
    double e = x * inv - 0x1p-19;      // below floor(ab/p) + 1 - 2^-20, above floor(ab/p) - 1
    uint32_t q = (uint32_t)fmax(e, 0.0);
    uint32_t r = a * b - q * p;  // true value in [0, 2p), exact mod 2^32
    return r >= p ? r - p : r;
}

// a * w mod p for a < 2^32 and w < p < 2^31, given w's Shoup companion
// wq = floor(w 2^32 / p): q = floor(a wq / 2^32) is within 1 below
// floor(a w / p), so a w - q p lies in [0, 2p) and is exact mod 2^32.
__device__ __forceinline__ uint32_t shoup_mul(uint32_t a, uint32_t w, uint32_t wq, uint32_t p) {
    uint32_t q = __umulhi(a, wq);
    uint32_t r = a * w - q * p;
    return r >= p ? r - p : r;
}

// The Shoup companion floor(w 2^32 / p) of w < p < 2^31.  The fp64 estimate
// w 2^32 inv - 1 lies below the true quotient and within 2 of it (relative
// error 2^-52 on a value below 2^32), so at most two corrections follow.
__device__ __forceinline__ uint32_t shoup_companion(uint32_t w, uint32_t p, double inv) {
    uint64_t q = (uint64_t)fmax((double)w * 0x1p32 * inv - 1.0, 0.0);
    uint64_t r = ((uint64_t)w << 32) - q * p;
    while (r >= p) r -= p, q++;
    return (uint32_t)q;
}

__device__ __forceinline__ uint32_t addmod(uint32_t a, uint32_t b, uint32_t p) {
    uint32_t s = a + b;
    return s >= p ? s - p : s;
}

__device__ __forceinline__ uint32_t submod(uint32_t a, uint32_t b, uint32_t p) {
    return a >= b ? a - b : a + (p - b);
}

// ---------------------------------------------------------------------------
// host helpers
static inline uint32_t h_mulmod(uint32_t a, uint32_t b, uint32_t p) { return (uint32_t)(((uint64_t)a * b) % p); }
// Products mod a fixed p < 2^31 by Barrett reduction: m = floor((2^64 - 1) / p)
// gives a quotient estimate at most two below the true one, so no division.
struct HostMod {
    uint32_t p;
    uint64_t m;
    explicit HostMod(uint32_t p_) : p(p_), m(~0ull / p_) {}
    uint32_t mul(uint32_t a, uint32_t b) const {
        const uint64_t x = (uint64_t)a * b;
        uint64_t r = x - (uint64_t)(((unsigned __int128)x * m) >> 64) * p;
        while (r >= p) r -= p;
        return (uint32_t)r;
    }
};
static inline uint32_t h_powmod(uint32_t a, uint64_t e, uint32_t p) {
    uint64_t r = 1, x = a % p;
    while (e) {
        if (e & 1) r = r * x % p;
        x = x * x % p;
        e >>= 1;
    }
    return (uint32_t)r;
}
static inline uint32_t h_invmod(uint32_t a, uint32_t p) { return h_powmod(a, p - 2, p); }
static inline bool h_is_prime(uint32_t p) {
    if (p < 2) return false;
    for (uint32_t d = 2; (uint64_t)d * d <= p; d++)
        if (p % d == 0) return false;
    return true;
}

// Streams of the library.  Every stream the library creates goes through
// these, so the device memory cache below knows every stream that can still
// be using a freed block; stream_destroy waits for the stream first.
hipError_t stream_create(hipStream_t *st);
void stream_destroy(hipStream_t st);

// Device memory through a per-device cache of freed blocks (size classes
// within 12.5 % of the request): hipMalloc / hipFree cost tens of
// microseconds each and hipFree waits for the device, which a batch of
// hundreds of matrices with dozens of buffers each pays thousands of times.
// A freed block may still be in use by kernels in flight.  The cache hands it
// out again once marker events recorded on every library stream after the
// free have completed; until then a request of its size gets a fresh block,
// so no allocation waits for the device.  When hipMalloc fails the cache
// waits for the device, releases its blocks and retries.
void *dev_alloc(size_t bytes, size_t *cls, int *device, hipError_t *err);
void dev_free(void *ptr, size_t cls, int device);
void dev_cache_release(int device);  // waits for the device, frees its cached blocks

// Owned device allocation.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0, cls = 0;
    int device = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : ptr(o.ptr), bytes(o.bytes), cls(o.cls), device(o.device) { o.ptr = nullptr; }
    ~DevBuf() {
        if (ptr) dev_free(ptr, cls, device);
    }
    rr_status alloc(size_t n) {
        if (ptr) dev_free(ptr, cls, device);
        ptr = nullptr;
        bytes = n;
        hipError_t e = hipSuccess;
        ptr = dev_alloc(n == 0 ? 4 : n, &cls, &device, &e);
        if (!ptr) {
            rr::set_error("hipMalloc(%zu) failed: %s", n, hipGetErrorString(e));
            return RR_ENOMEM;
        }
        return RR_OK;
    }
    template <typename T>
    T *as() const { return (T *)ptr; }
};

}  // namespace rr

#define RR_HIP(call)                                                                               \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) {                                                                    \
            rr::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
            return RR_EDEVICE;                                                                     \
        }                                                                                          \
    } while (0)

#define RR_CHECK(cond, code, ...)          \
    do {                                   \
        if (!(cond)) {                     \
            rr::set_error(__VA_ARGS__);    \
            return code;                   \
        }                                  \
    } while (0)
