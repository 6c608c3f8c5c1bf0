"""gcs_sort_pairs_noalloc, the allocation-free sort, against gcs_sort_pairs.

Through the C ABI on device buffers the test allocates with the HIP runtime:
three calls in a row on the same scratch (the largest size first, then
smaller ones), an in-place call, and the size limit.  Bar: keys and
permutation bit-exact with gcs_sort_pairs (a stable sort) on the same keys."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    for f in (h.hipMalloc, h.hipMemcpy, h.hipFree, h.hipDeviceSynchronize):
        f.restype = ctypes.c_int
    return h


def keys_case(n, seed):
    r = np.random.default_rng(seed)
    k = r.integers(0, 1 << 64, n, dtype=np.uint64)      # top bit used too
    k[: n // 4] %= np.uint64(1000)                       # many duplicates
    k[n // 4: n // 3] = np.uint64(0x7FFFFFFFFFFFFFFF)
    return k


def test_noalloc_sort_matches_sort_pairs(gcs):
    H = hip()
    L = gcs.lib()
    cap = gcs.NOALLOC_MAX_KEYS
    bufs = [ctypes.c_void_p() for _ in range(3)]
    for b, size in zip(bufs, (8 * cap, 8 * cap, 4 * cap)):
        assert H.hipMalloc(ctypes.byref(b), size) == 0
    dk, dko, dp = (b.value for b in bufs)
    try:
        for call, n in enumerate((cap, 300_001, 1)):
            keys = keys_case(n, 11 + call)
            want_k, want_p = gcs.sort_pairs(keys)
            assert H.hipMemcpy(dk, keys.ctypes.data, 8 * n, H2D) == 0
            print("noalloc sort call %d: n=%d" % (call, n), flush=True)
            assert L.gcs_sort_pairs_noalloc(dk, n, dko, dp, None) == 0
            assert H.hipDeviceSynchronize() == 0
            got_k = np.zeros(n, np.uint64)
            got_p = np.zeros(n, np.int32)
            assert H.hipMemcpy(got_k.ctypes.data, dko, 8 * n, D2H) == 0
            assert H.hipMemcpy(got_p.ctypes.data, dp, 4 * n, D2H) == 0
            assert np.array_equal(got_k, want_k), call
            assert np.array_equal(got_p, want_p), call
        # In place: keys_out == keys.
        n = 70_000
        keys = keys_case(n, 5)
        want_k, want_p = gcs.sort_pairs(keys)
        assert H.hipMemcpy(dk, keys.ctypes.data, 8 * n, H2D) == 0
        print("noalloc sort in place: n=%d" % n, flush=True)
        assert L.gcs_sort_pairs_noalloc(dk, n, dk, dp, None) == 0
        assert H.hipDeviceSynchronize() == 0
        got_k = np.zeros(n, np.uint64)
        got_p = np.zeros(n, np.int32)
        assert H.hipMemcpy(got_k.ctypes.data, dk, 8 * n, D2H) == 0
        assert H.hipMemcpy(got_p.ctypes.data, dp, 4 * n, D2H) == 0
        assert np.array_equal(got_k, want_k) and np.array_equal(got_p, want_p)
        # Beyond the limit: an error, nothing queued; empty: nothing to do.
        assert L.gcs_sort_pairs_noalloc(dk, cap + 1, dko, dp, None) == gcs.ERR_CAPACITY
        assert b"at most" in L.gcs_last_error()
        assert L.gcs_sort_pairs_noalloc(dk, 0, dko, dp, None) == 0
        assert L.gcs_sort_pairs_noalloc(None, 5, dko, dp, None) == gcs.ERR_INVALID
    finally:
        for b in bufs:
            H.hipFree(b)
