"""tensorfim_amd -- MI355X-native exact maximal-frequent-itemset engine.

Python side of the package: a ctypes binding of the C ABI in include/tfim.h
(libtfim.so, hand-written gfx950 HIP kernels) and helpers that drive the C++
engine binary (`tfim_mine`, the reference CLI `exe <file> <thr>`).

The binding fails loudly if the native library is missing: there is no CPU
fallback anywhere in the product path.
"""
import ctypes
import os
import re
import subprocess

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "libtfim.so")
BIN_DIR = os.path.join(_PKG, "bin")

TFIM_OK = 0
PATH_AUTO, PATH_PACKED, PATH_MFMA_I8, PATH_MFMA_FP4 = 0, 1, 2, 3
_PATHS = {"auto": PATH_AUTO, "packed": PATH_PACKED, "mfma": PATH_MFMA_I8, "mfma_i8": PATH_MFMA_I8,
          "fp4": PATH_MFMA_FP4, "mfma_fp4": PATH_MFMA_FP4}

_lib = None


class TfimError(RuntimeError):
    pass


def _p(t):
    return ctypes.POINTER(t)


_SIGS = {
    "tfim_open": (ctypes.c_int, [ctypes.c_int, _p(ctypes.c_void_p)]),
    "tfim_close": (None, [ctypes.c_void_p]),
    "tfim_open_shared": (ctypes.c_int, [ctypes.c_void_p, _p(ctypes.c_void_p)]),
    "tfim_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "tfim_device_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, _p(ctypes.c_int32),
                                         _p(ctypes.c_int64), _p(ctypes.c_int32)]),
    "tfim_load_bitmaps": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64,
                                          ctypes.c_int64, ctypes.c_int32]),
    "tfim_load_bitmaps_replicated": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                                     ctypes.c_int64, ctypes.c_int64, ctypes.c_int32,
                                                     ctypes.c_int32]),
    "tfim_item_supports": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "tfim_count_node": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]),
    "tfim_count_node_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]),
    "tfim_count_nodes": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32]),
    "tfim_count_nodes_ex": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                            ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32]),
    "tfim_count_masks": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                         ctypes.c_void_p, ctypes.c_int32]),
    "tfim_config": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int64]),
    "tfim_wave_stats": (ctypes.c_int, [ctypes.c_void_p, _p(ctypes.c_int64), _p(ctypes.c_int64),
                                        _p(ctypes.c_int64), _p(ctypes.c_double)]),
    "tfim_wave_graph_stats": (ctypes.c_int, [ctypes.c_void_p, _p(ctypes.c_int64), _p(ctypes.c_int64)]),
    "tfim_wave_path_stats": (ctypes.c_int, [ctypes.c_void_p, _p(ctypes.c_int64), _p(ctypes.c_int64),
                                             _p(ctypes.c_int64), _p(ctypes.c_int64)]),
    "tfim_prep_begin": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                        ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p]),
    "tfim_prep_finish": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                         _p(ctypes.c_int64), _p(ctypes.c_int64), _p(ctypes.c_int64),
                                         _p(ctypes.c_int64), _p(ctypes.c_int32)]),
    "tfim_buf_alloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, _p(ctypes.c_void_p)]),
    "tfim_buf_free": (None, [ctypes.c_void_p, ctypes.c_void_p]),
    "tfim_buf_write": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                       ctypes.c_int64]),
    "tfim_buf_read": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                      ctypes.c_int64]),
    "tfim_pitch_words": (ctypes.c_int64, [ctypes.c_int64]),
    "tfim_fill_bernoulli": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                            ctypes.c_int64, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32]),
    "tfim_popcount_matrix": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32,
                                             ctypes.c_int32]),
    "tfim_popcount_matrix_plan": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                  ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32,
                                                  _p(ctypes.c_int32), _p(ctypes.c_int32)]),
    "tfim_time_popcount_matrix": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                  ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                                  ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                                  _p(ctypes.c_double)]),
    "tfim_time_count_node": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                             ctypes.c_int32, _p(ctypes.c_double)]),
    "tfim_sync": (ctypes.c_int, [ctypes.c_void_p]),
    "tfim_mlp_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, _p(ctypes.c_void_p)]),
    "tfim_mlp_destroy": (None, [ctypes.c_void_p]),
    "tfim_mlp_get_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "tfim_mlp_set_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "tfim_mlp_forward": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                         ctypes.c_int32]),
    "tfim_mlp_forward_submit": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    "tfim_mlp_forward_wait": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "tfim_mlp_gradients": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_int32, ctypes.c_void_p, _p(ctypes.c_float)]),
    "tfim_mlp_train_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_int32, ctypes.c_float, _p(ctypes.c_float)]),
    "tfim_mlp_train": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_int32, ctypes.c_float, ctypes.c_int32, _p(ctypes.c_float)]),
}

EXPORTED_SYMBOLS = sorted(_SIGS)


def load_library(path=LIB_PATH):
    """Load libtfim.so; raises TfimError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise TfimError("libtfim.so not built at %s (run python -m tensorfim_amd.build)" % path)
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def pitch_words(n_bits):
    return int(load_library().tfim_pitch_words(int(n_bits)))


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class DeviceBuffer:
    """Device memory owned by the library (tfim_buf)."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        h = ctypes.c_void_p()
        ctx._check(ctx._lib.tfim_buf_alloc(ctx._h, self.nbytes, ctypes.byref(h)))
        self._h = h

    def write(self, arr, offset=0):
        a = np.ascontiguousarray(arr)
        self.ctx._check(self.ctx._lib.tfim_buf_write(self.ctx._h, self._h, int(offset), _ptr(a), a.nbytes))

    def read(self, dtype, count, offset=0):
        out = np.empty(int(count), dtype=dtype)
        self.ctx._check(self.ctx._lib.tfim_buf_read(self.ctx._h, self._h, int(offset), _ptr(out), out.nbytes))
        return out

    def free(self):
        if self._h:
            self.ctx._lib.tfim_buf_free(self.ctx._h, self._h)
            self._h = None


class Context:
    """One tfim_ctx: a device, a stream, resident item bitmaps."""

    def __init__(self, device=0):
        self._lib = load_library()
        h = ctypes.c_void_p()
        rc = self._lib.tfim_open(int(device), ctypes.byref(h))
        if rc != TFIM_OK:
            raise TfimError("tfim_open(%d) failed with status %d" % (device, rc))
        self._h = h
        self.n_cols = 0
        self.n_items = 0

    def close(self):
        if self._h:
            self._lib.tfim_close(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != TFIM_OK:
            raise TfimError("tfim status %d: %s" % (rc, self._lib.tfim_last_error(self._h).decode()))

    def device_info(self):
        name = ctypes.create_string_buffer(256)
        cus, hbm, clk = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int32()
        self._check(self._lib.tfim_device_info(self._h, name, 256, ctypes.byref(cus), ctypes.byref(hbm),
                                               ctypes.byref(clk)))
        return {"name": name.value.decode(), "compute_units": cus.value, "hbm_bytes": hbm.value,
                "clock_mhz": clk.value}

    def load_bitmaps(self, rows, n_bits, n_cols=None, replicate=1):
        """rows: uint64 array [n_items, words]; the first n_cols rows are counted."""
        rows = np.ascontiguousarray(rows, dtype=np.uint64)
        n_items, words = rows.shape
        n_cols = n_items if n_cols is None else int(n_cols)
        if replicate == 1:
            self._check(self._lib.tfim_load_bitmaps(self._h, _ptr(rows), n_items, int(n_bits), words, n_cols))
        else:
            self._check(self._lib.tfim_load_bitmaps_replicated(self._h, _ptr(rows), n_items, int(n_bits), words,
                                                               n_cols, int(replicate)))
        self.n_items, self.n_cols = n_items, n_cols

    def item_supports(self):
        out = np.empty(self.n_items, dtype=np.int32)
        self._check(self._lib.tfim_item_supports(self._h, _ptr(out)))
        return out

    def count_node(self, items):
        """Plain path: supports of items+{e} for every survivor e (getResFrequencyCuda)."""
        it = np.ascontiguousarray(items, dtype=np.int32)
        out = np.empty(self.n_cols, dtype=np.int32)
        self._check(self._lib.tfim_count_node(self._h, _ptr(it), len(it), _ptr(out)))
        return out

    def count_node_noalloc(self, items, out_dev_ptr):
        """tfim_count_node_noalloc: the plain path's counts written to the
        caller's device buffer at `out_dev_ptr` (n_cols int32), asynchronous
        on the context's stream, without allocating; call sync() before
        reading the buffer."""
        it = np.ascontiguousarray(items, dtype=np.int32)
        self._check(self._lib.tfim_count_node_noalloc(self._h, _ptr(it), len(it), ctypes.c_void_p(int(out_dev_ptr))))

    def time_count_node(self, items, warmup=3, reps=100):
        """Mean kernel ms of the plain per-node count for one node (HIP events)."""
        it = np.ascontiguousarray(items, dtype=np.int32)
        ms = ctypes.c_double()
        self._check(self._lib.tfim_time_count_node(self._h, _ptr(it), len(it), int(warmup), int(reps),
                                                   ctypes.byref(ms)))
        return ms.value

    def count_nodes(self, prefixes, tails=None, path="auto"):
        """Batched path.  prefixes: list of item lists; tails: None (all
        survivors) or list of column lists.  Returns a list of int32 arrays."""
        item_off = np.zeros(len(prefixes) + 1, dtype=np.int32)
        item_off[1:] = np.cumsum([len(p) for p in prefixes])
        item_ids = np.ascontiguousarray(np.concatenate([np.asarray(p, dtype=np.int32) for p in prefixes])
                                        if prefixes else np.zeros(0, np.int32), dtype=np.int32)
        n = len(prefixes)
        if tails is None:
            out = np.empty(max(1, n * self.n_cols), dtype=np.int32)
            self._check(self._lib.tfim_count_nodes(self._h, _ptr(item_ids), _ptr(item_off), n, None, None,
                                                   _ptr(out), _PATHS[path]))
            return [out[i * self.n_cols:(i + 1) * self.n_cols] for i in range(n)]
        col_off = np.zeros(n + 1, dtype=np.int32)
        col_off[1:] = np.cumsum([len(t) for t in tails])
        col_ids = np.ascontiguousarray(np.concatenate([np.asarray(t, dtype=np.int32) for t in tails])
                                       if n else np.zeros(0, np.int32), dtype=np.int32)
        out = np.empty(max(1, int(col_off[-1])), dtype=np.int32)
        self._check(self._lib.tfim_count_nodes(self._h, _ptr(item_ids), _ptr(item_off), n,
                                               _ptr(col_ids) if len(col_ids) else _ptr(np.zeros(1, np.int32)),
                                               _ptr(col_off), _ptr(out), _PATHS[path]))
        return [out[col_off[i]:col_off[i + 1]] for i in range(n)]

    def count_nodes_ex(self, prefixes, tails, excl, min_count, path="auto"):
        """Batched path with complement rows.  tails[i] is a column list, or
        None for a complement row whose exclusion set is excl[i] (survivor
        ids).  Returns, per row, its counts (explicit rows) or the pair
        (prefix support, number of non-excluded survivors other than the
        first prefix item with count >= min_count)."""
        n = len(prefixes)
        item_off = np.zeros(n + 1, dtype=np.int32)
        item_off[1:] = np.cumsum([len(p) for p in prefixes])
        item_ids = np.ascontiguousarray(np.concatenate([np.asarray(p, dtype=np.int32) for p in prefixes])
                                        if n else np.zeros(0, np.int32), dtype=np.int32)
        col_off = np.zeros(n + 1, dtype=np.int32)
        col_off[1:] = np.cumsum([0 if t is None else len(t) for t in tails])
        cols = [np.asarray(t, dtype=np.int32) for t in tails if t is not None]
        col_ids = np.ascontiguousarray(np.concatenate(cols) if cols else np.zeros(0, np.int32), dtype=np.int32)
        crows = np.asarray([i for i, t in enumerate(tails) if t is None], dtype=np.int32)
        words = (self.n_cols + 63) // 64
        ex = np.zeros((max(1, len(crows)), words), dtype=np.uint64)
        for j, i in enumerate(crows):
            for e in excl[i]:
                ex[j, e >> 6] |= np.uint64(1) << np.uint64(e & 63)
        n_exp = int(col_off[-1])
        out = np.empty(max(1, n_exp + 2 * len(crows)), dtype=np.int32)
        one = np.zeros(1, np.int32)
        self._check(self._lib.tfim_count_nodes_ex(
            self._h, _ptr(item_ids), _ptr(item_off), n, _ptr(col_ids) if len(col_ids) else None, _ptr(col_off),
            _ptr(crows) if len(crows) else _ptr(one), len(crows), _ptr(ex), int(min_count), _ptr(out),
            _PATHS[path]))
        res, j = [], 0
        for i, t in enumerate(tails):
            if t is None:
                res.append(out[n_exp + 2 * j:n_exp + 2 * j + 2])
                j += 1
            else:
                res.append(out[col_off[i]:col_off[i + 1]])
        return res

    def count_masks(self, prefixes, columns, path="auto"):
        """Batched path with rows as bit masks (at most 64 dense ids loaded):
        prefixes and columns are lists of id lists; row i's counts come back
        for its columns in ascending id order."""
        n = len(prefixes)
        pm = np.zeros(max(1, n), dtype=np.uint64)
        cm = np.zeros(max(1, n), dtype=np.uint64)
        for i in range(n):
            for x in prefixes[i]:
                pm[i] |= np.uint64(1) << np.uint64(x)
            for x in columns[i]:
                cm[i] |= np.uint64(1) << np.uint64(x)
        cnt = [bin(int(cm[i])).count("1") for i in range(n)]
        off = np.zeros(n + 1, dtype=np.int64)
        off[1:] = np.cumsum(cnt)
        out = np.empty(max(1, int(off[-1])), dtype=np.int32)
        self._check(self._lib.tfim_count_masks(self._h, _ptr(pm), _ptr(cm), n, _ptr(out), _PATHS[path]))
        return [out[off[i]:off[i + 1]] for i in range(n)]

    def config(self, key, value):
        self._check(self._lib.tfim_config(self._h, key.encode(), int(value)))

    def wave_stats(self):
        a, b, c, d = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_double()
        self._check(self._lib.tfim_wave_stats(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                              ctypes.byref(d)))
        g, r = ctypes.c_int64(), ctypes.c_int64()
        self._check(self._lib.tfim_wave_graph_stats(self._h, ctypes.byref(g), ctypes.byref(r)))
        l2, lds, mf, fu = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        self._check(self._lib.tfim_wave_path_stats(self._h, ctypes.byref(l2), ctypes.byref(lds), ctypes.byref(mf),
                                                   ctypes.byref(fu)))
        return {"launches": a.value, "lds_launches": b.value, "rows": c.value, "kernel_ms": d.value,
                "graphs": g.value, "graph_launches": r.value, "l2_launches": l2.value, "mfma_launches": mf.value,
                "fused_launches": fu.value}

    def buffer(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def fill_bernoulli(self, buf, rows, n_bits, pitch, seed, matrix, p):
        thr32 = int(np.floor(p * 2.0 ** 32))
        self._check(self._lib.tfim_fill_bernoulli(self._h, buf._h, int(rows), int(n_bits), int(pitch),
                                                  ctypes.c_uint64(seed), int(matrix), thr32))

    def popcount_matrix(self, A, B, C, b, d, pitch, path="auto", split_k=1):
        self._check(self._lib.tfim_popcount_matrix(self._h, A._h, B._h, C._h, int(b), int(d), int(pitch),
                                                   _PATHS[path], int(split_k)))

    def popcount_matrix_plan(self, A, B, C, b, d, pitch, path="auto"):
        """(tile, split_k) the tuner picks for this shape (split_k=0 calls)."""
        t, k = ctypes.c_int32(), ctypes.c_int32()
        self._check(self._lib.tfim_popcount_matrix_plan(self._h, A._h, B._h, C._h, int(b), int(d), int(pitch),
                                                        _PATHS[path], ctypes.byref(t), ctypes.byref(k)))
        return t.value, k.value

    def time_popcount_matrix(self, A, B, C, b, d, pitch, path="auto", split_k=1, warmup=1, reps=3):
        ms = ctypes.c_double()
        self._check(self._lib.tfim_time_popcount_matrix(self._h, A._h, B._h, C._h, int(b), int(d), int(pitch),
                                                        _PATHS[path], int(split_k), int(warmup), int(reps),
                                                        ctypes.byref(ms)))
        return ms.value

    def sync(self):
        self._check(self._lib.tfim_sync(self._h))


MLP_INPUTS, MLP_OUTPUTS, MLP_PARAMS = 14, 2, 5250


class Mlp:
    """The search policy's MLP (14 -> 64 -> 64 -> 2) on the device."""

    def __init__(self, ctx, seed=0):
        self.ctx = ctx
        h = ctypes.c_void_p()
        ctx._check(ctx._lib.tfim_mlp_create(ctx._h, ctypes.c_uint64(seed), ctypes.byref(h)))
        self._h = h
        self._pending_rows = 0

    def close(self):
        if self._h:
            self.ctx._lib.tfim_mlp_destroy(self._h)
            self._h = None

    def params(self):
        p = np.empty(MLP_PARAMS, dtype=np.float32)
        self.ctx._check(self.ctx._lib.tfim_mlp_get_params(self._h, _ptr(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self.ctx._check(self.ctx._lib.tfim_mlp_set_params(self._h, _ptr(p)))

    def forward(self, x, bf16=False):
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.empty((x.shape[0], MLP_OUTPUTS), dtype=np.float32)
        self.ctx._check(self.ctx._lib.tfim_mlp_forward(self._h, _ptr(x), x.shape[0], _ptr(out), int(bf16)))
        return out

    def forward_submit(self, x, bf16=False):
        x = np.ascontiguousarray(x, dtype=np.float32)
        self.ctx._check(self.ctx._lib.tfim_mlp_forward_submit(self._h, _ptr(x), x.shape[0], int(bf16)))
        self._pending_rows = x.shape[0]

    def forward_wait(self):
        out = np.empty((self._pending_rows, MLP_OUTPUTS), dtype=np.float32)
        self.ctx._check(self.ctx._lib.tfim_mlp_forward_wait(self._h, _ptr(out)))
        return out

    def gradients(self, x, y, m):
        x, y, m = (np.ascontiguousarray(a, dtype=np.float32) for a in (x, y, m))
        g = np.empty(MLP_PARAMS, dtype=np.float32)
        loss = ctypes.c_float()
        self.ctx._check(self.ctx._lib.tfim_mlp_gradients(self._h, _ptr(x), _ptr(y), _ptr(m), x.shape[0], _ptr(g),
                                                         ctypes.byref(loss)))
        return g, loss.value

    def train_step(self, x, y, m, lr=1e-3):
        x, y, m = (np.ascontiguousarray(a, dtype=np.float32) for a in (x, y, m))
        loss = ctypes.c_float()
        self.ctx._check(self.ctx._lib.tfim_mlp_train_step(self._h, _ptr(x), _ptr(y), _ptr(m), x.shape[0],
                                                          ctypes.c_float(lr), ctypes.byref(loss)))
        return loss.value

    def train(self, x, y, m, lr=1e-3, steps=1):
        """`steps` Adam steps on one batch staged once; the last step's loss."""
        x, y, m = (np.ascontiguousarray(a, dtype=np.float32) for a in (x, y, m))
        loss = ctypes.c_float()
        self.ctx._check(self.ctx._lib.tfim_mlp_train(self._h, _ptr(x), _ptr(y), _ptr(m), x.shape[0],
                                                     ctypes.c_float(lr), int(steps), ctypes.byref(loss)))
        return loss.value


# ---------------------------------------------------------------------------
# Engine binary (reference CLI) and Results files.

def binary(name):
    p = os.path.join(BIN_DIR, name)
    if not os.path.exists(p):
        raise TfimError("%s not built (run python -m tensorfim_amd.build)" % p)
    return p


def results_path(trans_file, support):
    """<file>-<%f>=Results.txt, as kernel_bmma.cu:1291-1292 names it."""
    return "%s-%f=Results.txt" % (trans_file, support)


def mine(trans_file, support, engine="mafia", out=None, extra=(), timeout=None):
    """Run the C++ engine; returns (results path, parsed stdout key/values)."""
    cmd = [binary("tfim_mine"), trans_file, repr(float(support)), "--engine", engine]
    if out:
        cmd += ["--out", out]
    cmd += list(extra)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise TfimError("tfim_mine failed (%d):\n%s\n%s" % (r.returncode, r.stdout[-4000:], r.stderr[-4000:]))
    stats = {}
    for ln in r.stdout.splitlines():
        if ":" in ln:
            k, v = ln.split(":", 1)
            stats[k.strip()] = v.strip()
    return (out or results_path(trans_file, support)), stats


_LINE = re.compile(r"^\d+th:\s+\d+\s+\{\s*([^}]*)\}\s+support:\s+\S+\s+frequency:\s+(\d+)")


def load_results(path):
    """{frozenset(items): frequency}, the comparison of scripts/verify_all.py:16-23."""
    sets = {}
    with open(path, encoding="utf-8", errors="replace") as f:
        for ln in f:
            m = _LINE.match(ln.strip())
            if m:
                sets[frozenset(int(x) for x in m.group(1).split())] = int(m.group(2))
    return sets


def load_header(path):
    """Header fields of a Results file (name -> string value)."""
    h = {}
    with open(path, encoding="utf-8", errors="replace") as f:
        for ln in f:
            if ln.startswith("=====") and "List of MFIs" in ln:
                break
            if ":" in ln:
                k, v = ln.split(":", 1)
                h.setdefault(k.strip(), v.strip())
    return h
