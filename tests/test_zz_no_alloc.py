"""lus_counting_sort_noalloc (static device scratch, no allocation) against
lus_counting_sort (the entry point that stages its own device buffers): the
same permutation and cell-start table, bit for bit, on two calls in a row, the
second with a smaller input.  Device buffers for the inputs and outputs come
from the HIP runtime the library itself links (libamdhip64.so.7)."""
import ctypes as C

import numpy as np
import pytest

from lustrine_amd import gpu as G
from lustrine_amd import scenes

pytestmark = pytest.mark.gpu


class Hip:
    def __init__(self):
        G.load()
        self.lib = C.CDLL("libamdhip64.so.7")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipDeviceSynchronize.argtypes = []
        self.lib.hipSetDevice.argtypes = [C.c_int]
        assert self.lib.hipSetDevice(0) == 0
        self.ptrs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), max(nbytes, 4)) == 0
        self.ptrs.append(p.value)
        return p.value

    def upload(self, a):
        a = np.ascontiguousarray(a)
        d = self.alloc(a.nbytes)
        assert self.lib.hipMemcpy(d, a.ctypes.data, a.nbytes, 1) == 0
        return d

    def download(self, d, n, dtype=np.int32):
        out = np.zeros(n, dtype)
        assert self.lib.hipMemcpy(out.ctypes.data, d, out.nbytes, 2) == 0
        return out

    def close(self):
        self.lib.hipDeviceSynchronize()
        for p in self.ptrs:
            self.lib.hipFree(p)
        self.ptrs = []


def glibc_rand_keys(num, mod):
    libc = C.CDLL("libc.so.6")
    libc.srand(1)
    return np.array([libc.rand() % mod for _ in range(num)], np.int32)


def _noalloc_sort(hip, keys, num_cells):
    n = keys.shape[0]
    d_keys = hip.upload(keys)
    d_counts = hip.alloc(4 * (num_cells + 1))
    d_sorted = hip.alloc(4 * n)
    G.counting_sort_noalloc(d_keys, n, num_cells, d_counts, d_sorted, None)
    assert hip.lib.hipDeviceSynchronize() == 0
    return hip.download(d_sorted, n), hip.download(d_counts, num_cells + 1)


def test_noalloc_sort_matches_staged_sort_twice():
    hip = Hip()
    try:
        # call 1: the cell keys of the 1,048,576-particle dam break's lattice
        cfg = scenes.cfg2()
        g = G.grid(cfg.cell_size, *cfg.grid)
        keys1 = G.cell_ids(g, cfg.positions())
        C1 = cfg.grid[0] * cfg.grid[1] * cfg.grid[2]
        perm1, counts1 = _noalloc_sort(hip, keys1, C1)
        want_perm1, want_counts1 = G.counting_sort(keys1, C1)
        assert np.array_equal(perm1, want_perm1) and np.array_equal(counts1, want_counts1)
        # call 2, smaller: the reference's unit-test keys (20,000 in [0, 1987))
        keys2 = glibc_rand_keys(20000, 1987)
        perm2, counts2 = _noalloc_sort(hip, keys2, 1987)
        want_perm2, want_counts2 = G.counting_sort(keys2, 1987)
        assert np.array_equal(perm2, want_perm2) and np.array_equal(counts2, want_counts2)
        assert np.array_equal(keys2[perm2], np.sort(keys2))
    finally:
        hip.close()


def test_noalloc_sort_rejects_sizes_beyond_the_scratch():
    hip = Hip()
    try:
        # non-null placeholders: the sizes are checked before any pointer is used
        d = hip.alloc(64)
        big = G.load().lus_counting_sort_noalloc(d, (1 << 21) + 1, 16, d, d, None)
        assert big == 1 and "exceed the static scratch" in G.last_error()
        cells = G.load().lus_counting_sort_noalloc(d, 16, (1 << 22) + 1, d, d, None)
        assert cells == 1 and "exceed the static scratch" in G.last_error()
    finally:
        hip.close()
