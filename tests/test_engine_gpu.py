"""End-to-end engine parity on the GPU.

Every engine (plain per-node, wave, batched MAFIA with and without the learned
policy) must reproduce the reference's output set exactly: the gold Results
files for the six FIMI datasets in the checkout (scripts/verify_all.py
comparison: {frozenset(items): frequency}), and the oracle's reference DFS on
the synthetic configurations.
"""
import os
import subprocess

import pytest

import oracle_lib
import tensorfim_amd as tf
from conftest import GOLD_PAIRS, gold_path

pytestmark = pytest.mark.gpu

ENGINES = [("plain", []), ("wave", []), ("mafia", []), ("mafia", ["--learned"]), ("mafia", ["--compl", "on"]),
           ("mafia", ["--compl", "off"]), ("mafia", ["--spec", "off"]), ("mafia", ["--rows", "lists"]),
           ("mafia", ["--graphs", "1"])]
HEADER_KEYS = ["transNum", "dimension", "itemIDmax", "dimensionReduced", "transNumReduced", "MFIsNumber",
               "transLengthMean", "transLengthReducedMean"]


def _id(e):
    return "-".join([e[0]] + [x.strip("-") for x in e[1]])


@pytest.mark.parametrize("engine", ENGINES, ids=_id)
@pytest.mark.parametrize("ds,thr", GOLD_PAIRS)
def test_gold_identity(fimi_dir, tmp_path, ds, thr, engine):
    name, extra = engine
    if name == "plain" and ds == "connect.txt" and thr == 0.85:
        # 142k nodes of single-node launches: the reference's worker pool
        extra = extra + ["--workers", "8"]
    out = str(tmp_path / ("%s-%s.res" % (ds, name)))
    path, stats = tf.mine(str(fimi_dir / ds), thr, engine=name, out=out, extra=extra, timeout=300)
    print(ds, thr, name, extra, {k: stats.get(k) for k in ("MFIsNumber", "rows", "waves", "computation_time(s)")},
          flush=True)
    got, gold = tf.load_results(path), tf.load_results(gold_path(ds, thr))
    assert got == gold
    hg, hr = tf.load_header(path), tf.load_header(gold_path(ds, thr))
    for k in HEADER_KEYS:
        assert hg[k] == hr[k], k


def _gen(cfg, path, rows=None):
    cmd = [tf.binary("tfim_gen"), str(cfg), str(path)]
    if rows:
        cmd += ["--rows", str(rows)]
    subprocess.run(cmd, check=True)


@pytest.mark.parametrize("cfg,thr,rows", [("dense-small", 0.6, None), ("gene", 0.12, None),
                                          ("dense-replicated", 0.9, 6000), ("clickstream", 0.02, 100000)])
def test_synthetic_configs_match_oracle(tmp_path, cfg, thr, rows):
    data = tmp_path / ("%s.txt" % cfg)
    _gen(cfg, data, rows)
    ref = tmp_path / "oracle.res"
    oracle_lib.mine(data, thr, threads=16, out=ref)
    exp = tf.load_results(str(ref))
    assert exp, "oracle found no MFI"
    for name, extra in ENGINES:
        if name == "plain" and cfg == "gene":
            continue   # 19k launches; the wave engine runs the same enumeration
        out = tmp_path / ("%s.res" % _id((name, extra)))
        path, stats = tf.mine(str(data), thr, engine=name, out=str(out), extra=extra, timeout=600)
        print(cfg, name, extra, {k: stats.get(k) for k in ("MFIsNumber", "rows", "computation_time(s)")}, flush=True)
        assert tf.load_results(path) == exp, (name, extra)


def test_replicated_dense_scales_exactly(tmp_path):
    """x k replication scales every support by k; thresholds 0.9*N are
    non-integers for both sizes, so the MFI set is unchanged."""
    data = tmp_path / "c4.txt"
    _gen("dense-replicated", data, 3001)
    base = tmp_path / "base.res"
    oracle_lib.mine(data, 0.9, threads=16, out=base)
    exp = {s: 16 * f for s, f in tf.load_results(str(base)).items()}
    path, _ = tf.mine(str(data), 0.9, engine="mafia", out=str(tmp_path / "rep.res"), extra=["--replicate", "16"])
    got = tf.load_results(path)
    # single-item MFIs report occurrence counts, which also scale by k
    assert got == exp


HEADER_ALL = HEADER_KEYS + ["itemIDmin", "transLengthMin", "transLengthMax", "transLengthReducedMin",
                            "transLengthReducedMax"]


@pytest.mark.parametrize("cfg,thr,rows", [("dense-small", 0.6, None), ("gene", 0.12, None),
                                          ("clickstream", 0.02, 100000), ("dense-replicated", 0.9, 6000)])
def test_device_preprocessing_matches_host(tmp_path, cfg, thr, rows):
    """Counts, reduction and bitmaps built on the device (LDS or global
    histogram by id range) give the host pass's header and MFI set."""
    data = tmp_path / ("%s.txt" % cfg)
    _gen(cfg, data, rows)
    got = {}
    for prep in ("gpu", "host"):
        path, st = tf.mine(str(data), thr, engine="mafia", out=str(tmp_path / (prep + ".res")), extra=["--prep", prep],
                           timeout=300)
        got[prep] = (tf.load_results(path), {k: tf.load_header(path)[k] for k in HEADER_ALL})
    assert got["gpu"] == got["host"]


def test_device_preprocessing_edge_rows(tmp_path):
    """Empty lines, a last line without newline, repeated ids (the engine then
    falls back to the reference enumeration), item 0, rows that lose every
    item: device and host preprocessing agree and match the oracle."""
    lines = ["1 2 3", "", "2 3 3", "0 1 2", "7", "", "0 2 3 1", "3 2", "9 8", "1 2 0"]
    data = tmp_path / "edge.txt"
    data.write_text("\n".join(lines))
    ref = tmp_path / "oracle.res"
    oracle_lib.mine(data, 0.3, threads=2, out=ref)
    for prep in ("gpu", "host"):
        path, st = tf.mine(str(data), 0.3, engine="mafia", out=str(tmp_path / (prep + ".res")), extra=["--prep", prep],
                           timeout=120)
        assert tf.load_results(path) == tf.load_results(str(ref)), prep
        hg, hr = tf.load_header(path), tf.load_header(str(ref))
        for k in HEADER_KEYS:
            assert hg[k] == hr[k], (prep, k)


def test_device_preprocessing_staged_upload():
    """Blocks of 64 MB and more go up through pinned buffers filled by several
    host threads (tfim_prep_begin): about 24 M ids (96 MB, per-thread shares
    that are not whole 8 MB pieces) reach the device intact.  Checked: the
    occurrence counts equal numpy's bincount, the kept row count equals the
    non-empty rows, and for a few items the counts against every item equal
    the rows' co-occurrence counts (a misplaced piece moves ids between
    rows)."""
    import ctypes

    import numpy as np
    import scipy.sparse as sp

    rng = np.random.default_rng(11)
    n_rows, id_max = 600_001, 1999
    lens = rng.integers(0, 81, n_rows)
    off = np.zeros(n_rows + 1, np.int64)
    np.cumsum(lens, out=off[1:])
    n_ids = int(off[-1])
    assert n_ids * 4 >= 64 << 20
    ids = rng.integers(0, id_max + 1, n_ids, dtype=np.int32)
    with tf.Context() as ctx:
        lib = ctx._lib
        freq = np.zeros(id_max + 1, np.int32)
        ctx._check(lib.tfim_prep_begin(ctx._h, ids.ctypes.data, n_ids, off.ctypes.data, n_rows, id_max,
                                       freq.ctypes.data))
        assert np.array_equal(freq, np.bincount(ids, minlength=id_max + 1))
        dense_of = np.arange(id_max + 1, dtype=np.int32)
        n_red, lmin, lmax, lsum, dup = (ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64(),
                                        ctypes.c_int32())
        ctx._check(lib.tfim_prep_finish(ctx._h, dense_of.ctypes.data, id_max + 1, id_max + 1, ctypes.byref(n_red),
                                        ctypes.byref(lmin), ctypes.byref(lmax), ctypes.byref(lsum),
                                        ctypes.byref(dup)))
        ctx.n_items = ctx.n_cols = id_max + 1
        assert n_red.value == int((lens > 0).sum())
        m = sp.csr_matrix((np.ones(n_ids, np.int32), ids, off), shape=(n_rows, id_max + 1))
        m.sum_duplicates()
        m.data[:] = 1
        mt = m.T.tocsr()
        for x in (0, 7, 1024, id_max):
            exp = np.asarray((mt[x] @ m).todense()).ravel()
            assert np.array_equal(ctx.count_node([x]), exp), x
