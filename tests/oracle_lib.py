"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so) for tests.

The oracle is the checker: tests compare the HIP library and engine against
it; nothing in tensorfim_amd/ imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB = os.path.join(ORACLE_DIR, "_build", "liboracle.so")
CLI = os.path.join(ORACLE_DIR, "_build", "tfim_oracle")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        vp = ctypes.c_void_p
        L.oracle_popcount_matrix.argtypes = [vp, vp, vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                             ctypes.c_int64, ctypes.c_int]
        L.oracle_popcount_matrix.restype = None
        L.oracle_bernoulli_row.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int64,
                                           ctypes.c_int64, ctypes.c_uint32, vp]
        L.oracle_bernoulli_row.restype = None
        L.oracle_proto_gen.argtypes = [vp, vp, vp]
        L.oracle_proto_gen.restype = None
        L.oracle_mlp_loss_grad.argtypes = [vp, vp, vp, vp, ctypes.c_int, vp, vp]
        L.oracle_mlp_loss_grad.restype = ctypes.c_double
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def popcount_matrix(A, B, threads=8):
    """C[i][j] = popcount(A_i & B_j) (bmma_bench.cu:180-191)."""
    A = np.ascontiguousarray(A, dtype=np.uint64)
    B = np.ascontiguousarray(B, dtype=np.uint64)
    assert A.shape[1] == B.shape[1]
    C = np.empty((A.shape[0], B.shape[0]), dtype=np.int32)
    lib().oracle_popcount_matrix(_p(A), _p(B), _p(C), A.shape[0], B.shape[0], A.shape[1], B.shape[0], threads)
    return C


def bernoulli_row(seed, matrix, row, n_bits, words, thr32):
    out = np.empty(words, dtype=np.uint64)
    lib().oracle_bernoulli_row(ctypes.c_uint64(seed), matrix, ctypes.c_uint64(row), n_bits, words, thr32, _p(out))
    return out


def node_counts(rows, prefix, n_cols):
    """Supports of prefix+{e} for e < n_cols: AND of the prefix rows, then counts."""
    mask = np.bitwise_and.reduce(rows[np.asarray(prefix)], axis=0)
    return popcount_matrix(mask[None, :], rows[:n_cols])[0]


def mine(path, support, threads=8, replicate=0, out=None, max_nodes=0, timeout=600):
    """Run the oracle CLI (reference DFS); returns parsed stdout stats."""
    if not os.path.exists(CLI):
        build()
    cmd = [CLI, str(path), repr(float(support)), "--threads", str(threads)]
    if replicate:
        cmd += ["--replicate", str(replicate)]
    if out:
        cmd += ["--out", str(out)]
    if max_nodes:
        cmd += ["--max-nodes", str(max_nodes)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, check=True)
    stats = {}
    for ln in r.stdout.splitlines():
        if ":" in ln:
            k, v = ln.split(":", 1)
            stats[k.strip()] = v.strip()
    return stats


def mlp_loss_grad(params, X, Y, M, want_grad=True):
    """Loss, logits[n][2] and gradients (float64) of the policy MLP."""
    P = np.ascontiguousarray(params, dtype=np.float64)
    X, Y, M = (np.ascontiguousarray(a, dtype=np.float32) for a in (X, Y, M))
    logits = np.empty((X.shape[0], 2), dtype=np.float64)
    g = np.empty(P.shape[0], dtype=np.float64) if want_grad else None
    loss = lib().oracle_mlp_loss_grad(_p(P), _p(X), _p(Y), _p(M), X.shape[0], _p(logits),
                                      _p(g) if want_grad else None)
    return loss, logits, g


def proto_bitmaps():
    """batch_proto.cu's synthetic bitmaps: [512][2^23/64] u64 plus patterns."""
    B = np.empty((512, (1 << 23) // 64), dtype=np.uint64)
    pat = np.zeros(80 * 6, dtype=np.int32)
    plen = np.zeros(80, dtype=np.int32)
    lib().oracle_proto_gen(_p(B), _p(pat), _p(plen))
    return B, [pat[6 * i:6 * i + plen[i]].tolist() for i in range(80)]
