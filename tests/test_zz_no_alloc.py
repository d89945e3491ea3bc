"""rr_ntt_noalloc (the NTT entry point that allocates no device memory)
against rr_ntt, through the C ABI, on device buffers the test owns: two calls
in a row, the second on a smaller input, forward and inverse, on the null
stream and on a stream of the caller's."""
import ctypes as C

import numpy as np
import pytest

P = 2013265921
pytestmark = pytest.mark.gpu
MAX_WORDS = 1 << 22


class _Hip:
    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.h.hipStreamDestroy.argtypes = [C.c_void_p]
        self.h.hipStreamSynchronize.argtypes = [C.c_void_p]

    def check(self, rc):
        assert rc == 0, f"HIP error {rc}"

    def malloc(self, nbytes):
        ptr = C.c_void_p()
        self.check(self.h.hipMalloc(C.byref(ptr), nbytes))
        return ptr

    def h2d(self, dst, a):
        self.check(self.h.hipMemcpy(dst, a.ctypes.data, a.nbytes, 1))

    def d2h(self, a, src):
        self.check(self.h.hipMemcpy(a.ctypes.data, src, a.nbytes, 2))


def _lib():
    from rust_rank_amd import engine
    assert engine.device_count() >= 1, "no GPU visible"
    L = engine.lib()
    L.rr_ntt_noalloc.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_int, C.c_void_p]
    L.rr_ntt_noalloc.restype = C.c_int
    return engine, L


def _noalloc(hip, L, a, inverse, stream):
    count, N = a.shape
    logn = N.bit_length() - 1
    buf = hip.malloc(a.nbytes)
    try:
        hip.h2d(buf, np.ascontiguousarray(a))
        rc = L.rr_ntt_noalloc(buf, logn, count, int(inverse), stream)
        assert rc == 0, L.rr_last_error().decode()
        hip.check(hip.h.hipStreamSynchronize(stream))
        out = np.empty_like(a)
        hip.d2h(out, buf)
    finally:
        hip.check(hip.h.hipFree(buf))
    return out


@pytest.mark.parametrize("on_stream", [False, True])
def test_noalloc_ntt_equals_rr_ntt(on_stream):
    eng, L = _lib()
    hip = _Hip()
    stream = C.c_void_p()
    if on_stream:
        hip.check(hip.h.hipStreamCreate(C.byref(stream)))
    try:
        rng = np.random.default_rng(77 + on_stream)
        # the largest input the work buffer takes (four-step size), then a smaller one
        for logn, count in [(18, MAX_WORDS >> 18), (12, 5), (15, 3)]:
            print(f"noalloc ntt logn={logn} count={count} stream={on_stream}", flush=True)
            a = rng.integers(0, P, (count, 1 << logn)).astype(np.uint32)
            f = _noalloc(hip, L, a, False, stream)
            assert np.array_equal(f, eng.ntt(a))
            back = _noalloc(hip, L, f, True, stream)
            assert np.array_equal(back, eng.ntt(f, inverse=True))
            assert np.array_equal(back, a)
    finally:
        if on_stream:
            hip.check(hip.h.hipStreamDestroy(stream))


def test_noalloc_ntt_rejects_oversize():
    eng, L = _lib()
    hip = _Hip()
    buf = hip.malloc(4)
    try:
        rc = L.rr_ntt_noalloc(buf, 20, 5, 0, None)  # 5 x 2^20 words > 2^22
        assert rc == -1 and b"work buffer" in L.rr_last_error()
        assert L.rr_ntt_noalloc(buf, 21, 1, 0, None) == -1  # beyond 2^20 points
    finally:
        hip.check(hip.h.hipFree(buf))
