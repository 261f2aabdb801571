"""GPU: tfim_count_node_noalloc (the plain per-node count without any device
allocation, its partial counts in the library's fixed scratch array) gives
exactly what tfim_count_node gives, on two calls in a row -- the second on a
smaller database, so counts left in the scratch array by the first call would
show -- and refuses sizes beyond its limit with a clear error.  The output
buffer is caller-owned device memory, allocated here through the HIP runtime
libtfim itself is linked against (torch ships a HIP runtime of its own, and
two runtimes in one process do not share devices)."""
import ctypes

import numpy as np
import pytest

import oracle_lib
import tensorfim_amd as tf

pytestmark = pytest.mark.gpu


class DeviceInts:
    """n int32 of device memory from libtfim's own HIP runtime."""
    H2D, D2H = 1, 2

    def __init__(self, n):
        tf.load_library()
        self.hip = ctypes.CDLL("libamdhip64.so.7")     # the instance libtfim.so already loaded
        self.n = n
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(4 * n)) == 0
        self.ptr = p.value

    def write(self, a):
        a = np.ascontiguousarray(a, dtype=np.int32)
        assert self.hip.hipMemcpy(ctypes.c_void_p(self.ptr), a.ctypes.data_as(ctypes.c_void_p),
                                  ctypes.c_size_t(a.nbytes), self.H2D) == 0

    def read(self):
        a = np.empty(self.n, dtype=np.int32)
        assert self.hip.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr),
                                  ctypes.c_size_t(a.nbytes), self.D2H) == 0
        return a

    def free(self):
        self.hip.hipFree(ctypes.c_void_p(self.ptr))


def rand_rows(rng, n_items, n_bits, density):
    words = (n_bits + 63) // 64
    bits = rng.random((n_items, words * 64)) < density
    bits[:, n_bits:] = False
    return np.packbits(bits, axis=1, bitorder="little").view(np.uint64).reshape(n_items, words)


def noalloc_counts(ctx, items, n_cols):
    out = DeviceInts(n_cols)
    try:
        out.write(np.full(n_cols, -7, dtype=np.int32))
        ctx.count_node_noalloc(items, out.ptr)
        ctx.sync()
        return out.read()
    finally:
        out.free()


def test_noalloc_matches_count_node_twice():
    with tf.Context(0) as ctx:
        rng = np.random.default_rng(31)
        cases = [(3000, 1_000_003, [17, 400, 2999, 5]),   # 8 copies of 3,072 counts
                 (70, 5_001, [3, 69])]                      # smaller: stale partial counts would show
        for n_cols, n_bits, prefix in cases:
            print("case: %d survivors, %d transactions, prefix %s" % (n_cols, n_bits, prefix), flush=True)
            rows = rand_rows(rng, n_cols + 1, n_bits, 0.5)    # one extra prefix-only row
            ctx.load_bitmaps(rows, n_bits, n_cols)
            for p in (prefix, prefix + [n_cols]):
                ref = ctx.count_node(p)
                got = noalloc_counts(ctx, p, n_cols)
                np.testing.assert_array_equal(got, ref, err_msg="prefix %s" % p)
                np.testing.assert_array_equal(got, oracle_lib.node_counts(rows, p, n_cols))


def test_noalloc_rejects_sizes_beyond_its_limit():
    with tf.Context(0) as ctx:
        rng = np.random.default_rng(2)
        rows = rand_rows(rng, 8193, 300, 0.5)
        ctx.load_bitmaps(rows, 300, 8193)              # one survivor above TFIM_NOALLOC_MAX_COLS
        out = DeviceInts(8193)
        try:
            with pytest.raises(tf.TfimError, match="scratch limit"):
                ctx.count_node_noalloc([1, 2], out.ptr)
            ctx.load_bitmaps(rows[:100], 300, 100)
            with pytest.raises(tf.TfimError, match="prefix items"):
                ctx.count_node_noalloc(list(range(61)), out.ptr)
            with pytest.raises(tf.TfimError, match="out of range"):
                ctx.count_node_noalloc([1, 100], out.ptr)
        finally:
            out.free()
