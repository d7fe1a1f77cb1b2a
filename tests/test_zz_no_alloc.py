"""GPU: kpax_repeat_expand_noalloc, the allocation-free form of the weighted
repeat-expansion (frontier scan + findInd + repeatInd, KPAX.cu:310-321 and
Planner.cu:52-80), checked against kpax_repeat_expand on two calls in a row,
the second smaller than the first, and on an empty input.  The test's own
device buffers come from the HIP runtime the library runs on (hipMalloc
through ctypes); the entry point allocates nothing."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime libkpax is linked with."""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so", mode=ctypes.RTLD_GLOBAL)
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.bufs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), max(nbytes, 4)) == 0
        self.bufs.append(p.value)
        return p.value

    def upload(self, a):
        d = self.alloc(a.nbytes)
        if a.nbytes:
            assert self.L.hipMemcpy(d, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return d

    def download(self, d, n, dtype=np.uint32):
        a = np.zeros(n, dtype)
        if n:
            assert self.L.hipMemcpy(a.ctypes.data, d, a.nbytes, 2) == 0  # device to host
        return a

    def sync(self):
        assert self.L.hipDeviceSynchronize() == 0

    def free(self):
        for b in self.bufs:
            self.L.hipFree(b)
        self.bufs = []


def test_repeat_expand_noalloc_matches_repeat_expand(kpax):
    rng = np.random.default_rng(31)
    with kpax.KPAX(kpax.config_workload(1)) as p:
        hip = Hip()
        try:
            for n in (300000, 5000, 0):
                w = rng.integers(0, 4, n).astype(np.uint32)
                if n:
                    w[-3:] = 5
                want, total = p.repeat_expand(w)
                cap = max(total, 1)
                d_w = hip.upload(w)
                d_out = hip.upload(np.full(cap, 0xFFFFFFFF, np.uint32))
                d_total = hip.upload(np.full(1, 0xFFFFFFFF, np.uint32))
                hip.sync()
                print(f"noalloc repeat-expand n={n} total={total}", flush=True)
                p.repeat_expand_noalloc(d_w, n, d_out, cap, d_total)
                hip.sync()
                assert int(hip.download(d_total, 1)[0]) == total
                assert np.array_equal(hip.download(d_out, total), want)
                assert np.array_equal(want, np.repeat(np.arange(n, dtype=np.uint32), w))
        finally:
            hip.free()


def test_repeat_expand_noalloc_rejects_too_many_items(kpax):
    with kpax.KPAX(kpax.config_workload(1)) as p:
        with pytest.raises(kpax.KpaxError, match="KPAX_NOALLOC_MAX_ITEMS"):
            p.repeat_expand_noalloc(16, (1 << 24) + 1, 16, 1, 16)
