"""GPU test of the allocation-free entry point (sdim_simulate_noalloc) through the C ABI:
its records equal those of sdim_submit on the same jobs, on two calls in a row into one
device buffer (the second with a smaller input), and buffers that are too small or
metadata beyond its per-GPU array are refused.  The output buffer is allocated by the
test through the HIP runtime the library itself is linked against (looked up through the
library's handle; the entry point allocates nothing)."""
import ctypes

import numpy as np
import pytest

from noalloc_helpers import check_noalloc_against_submit

pytestmark = pytest.mark.gpu

def test_noalloc_matches_submit():
    from sdim_amd.engine import Engine
    eng = Engine()
    H = eng.L                              # its dependencies resolve hipMalloc & co.
    H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    H.hipFree.argtypes = [ctypes.c_void_p]
    H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    H.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    bufs = []

    def alloc(nbytes):
        p = ctypes.c_void_p()
        assert H.hipMalloc(ctypes.byref(p), nbytes) == 0
        bufs.append(p)
        return p.value

    def to_host(ptr, nbytes):
        out = np.zeros(nbytes, np.uint8)
        assert H.hipMemcpy(out.ctypes.data, ptr, nbytes, 2) == 0     # device to host
        return out

    def fill(ptr, value, nbytes):
        assert H.hipMemset(ptr, value, nbytes) == 0

    print("[noalloc] two calls into one device buffer", flush=True)
    try:
        check_noalloc_against_submit(eng, np.random.default_rng(2027), alloc, to_host, fill)
    finally:
        for p in bufs:
            H.hipFree(p)
        eng.close()
