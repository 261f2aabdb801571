"""GPU parity of the counting kernels against the CPU oracle (bit-exact).

Shapes cover the edge cases the reference's own checks target
(bmma_bench.cu:326-358, RESULTS.md:45): sparse random data (never all-ones),
transaction counts off multiples of 64 and 2048, column counts off multiples
of 8/64/128, prefixes of depth 1..31, an extra prefix-only row (item 0,
kernel_bmma.cu:402-403), empty tails.
"""
import numpy as np
import pytest

import oracle_lib
import tensorfim_amd as tf

pytestmark = pytest.mark.gpu


def rand_rows(rng, n_items, n_bits, density=0.3):
    words = (n_bits + 63) // 64
    bits = rng.random((n_items, words * 64)) < density
    bits[:, n_bits:] = False
    return np.packbits(bits, axis=1, bitorder="little").view(np.uint64).reshape(n_items, words)


@pytest.fixture(scope="module")
def ctx():
    c = tf.Context(0)
    print("device:", c.device_info(), flush=True)
    yield c
    c.close()


@pytest.mark.parametrize("n_bits,n_items", [(70, 9), (3200, 20), (2048 * 5 + 17, 131), (1_000_003, 70)])
def test_count_node_matches_oracle(ctx, n_bits, n_items):
    rng = np.random.default_rng(n_bits)
    rows = rand_rows(rng, n_items, n_bits, 0.55)
    n_cols = n_items - 1          # last row: prefix-only (item 0 quirk)
    ctx.load_bitmaps(rows, n_bits, n_cols)
    for k in (1, 2, 3, 7):
        prefix = rng.choice(n_items, size=k, replace=False)
        got = ctx.count_node(prefix)
        exp = oracle_lib.node_counts(rows, prefix, n_cols)
        np.testing.assert_array_equal(got, exp, err_msg="k=%d prefix=%s" % (k, prefix))


@pytest.mark.parametrize("shape", [108, 116, 132, 164, 208, 216, 232, 264, 408, 416, 432, 464])
def test_count_node_launch_shapes(ctx, shape):
    """Every launch shape of the per-node kernel (words per lane x columns per
    block) gives the oracle's counts, on ragged sizes."""
    rng = np.random.default_rng(shape)
    n_bits, n_items = 300_017, 139
    rows = rand_rows(rng, n_items, n_bits, 0.5)
    ctx.load_bitmaps(rows, n_bits, n_items - 2)
    ctx.config("node_shape", shape)
    try:
        for k in (1, 4):
            prefix = rng.choice(n_items, size=k, replace=False)
            got = ctx.count_node(prefix)
            np.testing.assert_array_equal(got, oracle_lib.node_counts(rows, prefix, n_items - 2),
                                          err_msg="shape=%d k=%d" % (shape, k))
    finally:
        ctx.config("node_shape", 0)


def test_item_supports(ctx):
    rng = np.random.default_rng(5)
    rows = rand_rows(rng, 37, 100_001, 0.2)
    ctx.load_bitmaps(rows, 100_001, 30)
    exp = np.array([int(np.unpackbits(r.view(np.uint8)).sum()) for r in rows], dtype=np.int32)
    np.testing.assert_array_equal(ctx.item_supports(), exp)


@pytest.mark.parametrize("n_bits,n_items", [(1100, 300), (3200, 75), (65_537, 40), (2_000_000, 33)])
def test_count_nodes_full_columns(ctx, n_bits, n_items):
    rng = np.random.default_rng(n_items)
    rows = rand_rows(rng, n_items, n_bits, 0.7)
    n_cols = n_items
    ctx.load_bitmaps(rows, n_bits, n_cols)
    prefixes = [rng.choice(n_items, size=int(rng.integers(1, min(31, n_items) + 1)), replace=False)
                for _ in range(77)]
    got = ctx.count_nodes(prefixes)
    for p, g in zip(prefixes, got):
        np.testing.assert_array_equal(g, oracle_lib.node_counts(rows, p, n_cols))


@pytest.mark.parametrize("n_bits,n_items", [(3200, 75), (300_001, 61)])
def test_count_nodes_tails(ctx, n_bits, n_items):
    rng = np.random.default_rng(7 + n_items)
    rows = rand_rows(rng, n_items, n_bits, 0.6)
    n_cols = n_items - 1
    ctx.load_bitmaps(rows, n_bits, n_cols)
    prefixes, tails = [], []
    for i in range(150):
        k = int(rng.integers(1, 12))
        prefixes.append(rng.choice(n_items, size=k, replace=False))
        t = int(rng.integers(0, n_cols + 1)) if i % 10 else 0      # some empty tails
        tails.append(np.sort(rng.choice(n_cols, size=t, replace=False)))
    got = ctx.count_nodes(prefixes, tails)
    for p, t, g in zip(prefixes, tails, got):
        exp = oracle_lib.node_counts(rows, p, n_cols)[t]
        np.testing.assert_array_equal(g, exp)


def test_replicated_bitmaps(ctx):
    """make_replica.py semantics on the device: every support scales by k."""
    rng = np.random.default_rng(11)
    base_bits, k, n_items = 49_046 // 16, 5, 12
    rows = rand_rows(rng, n_items, base_bits, 0.5)
    # host-side replication, bit by bit
    bits = np.unpackbits(rows.view(np.uint8), axis=1, bitorder="little")[:, :base_bits]
    rep = np.tile(bits, (1, k))
    words = (base_bits * k + 63) // 64
    pad = np.zeros((n_items, words * 64), dtype=np.uint8)
    pad[:, :base_bits * k] = rep
    rep_rows = np.packbits(pad, axis=1, bitorder="little").view(np.uint64).reshape(n_items, words)
    ctx.load_bitmaps(rows, base_bits, n_items, replicate=k)
    got_rep = [ctx.count_node([i, j]) for i, j in [(0, 1), (3, 7), (11, 2)]]
    ctx.load_bitmaps(rep_rows, base_bits * k, n_items)
    for (i, j), g in zip([(0, 1), (3, 7), (11, 2)], got_rep):
        np.testing.assert_array_equal(g, ctx.count_node([i, j]))
        np.testing.assert_array_equal(g, k * oracle_lib.node_counts(rows, [i, j], n_items))


def test_fill_bernoulli_matches_oracle(ctx):
    n_bits, rows = 100_000 + 13, 9
    pitch = tf.pitch_words(n_bits)
    buf = ctx.buffer(rows * pitch * 8)
    thr32 = int(np.floor(0.3 * 2 ** 32))
    ctx.fill_bernoulli(buf, rows, n_bits, pitch, seed=1, matrix=0, p=0.3)
    got = buf.read(np.uint64, rows * pitch).reshape(rows, pitch)
    for r in range(rows):
        exp = oracle_lib.bernoulli_row(1, 0, r, n_bits, pitch, thr32)
        np.testing.assert_array_equal(got[r], exp)
    dens = np.unpackbits(got.view(np.uint8)).sum() / (rows * n_bits)
    assert abs(dens - 0.3) < 0.01
    buf.free()


@pytest.mark.parametrize("path", ["packed", "mfma", "mfma128", "mfma256", "fp4"])
@pytest.mark.parametrize("b,d,n_bits,split", [(130, 70, 4096 + 65, 1), (256, 256, 65_536, 1),
                                               (200, 513, 40_000, 3), (64, 33, 2048 * 9, 8),
                                               (1, 1, 64, 1), (515, 300, 2048 * 3 + 1, 2),
                                               (40, 24, (1 << 24) + 2048 * 3, 1)])
def test_popcount_matrix_matches_oracle(ctx, path, b, d, n_bits, split):
    rng = np.random.default_rng(b * 1000 + d)
    pitch = tf.pitch_words(n_bits)
    A = np.zeros((b, pitch), np.uint64)
    B = np.zeros((d, pitch), np.uint64)
    w = (n_bits + 63) // 64
    A[:, :w] = rand_rows(rng, b, n_bits, 0.3)
    B[:, :w] = rand_rows(rng, d, n_bits, 0.3)
    dA, dB, dC = ctx.buffer(A.nbytes), ctx.buffer(B.nbytes), ctx.buffer(b * d * 4)
    dA.write(A)
    dB.write(B)
    tile = {"mfma128": 128, "mfma256": 256}.get(path, 0)
    ctx.config("mfma_tile", tile)
    try:
        ctx.popcount_matrix(dA, dB, dC, b, d, pitch, path=path if path in ("packed", "fp4") else "mfma",
                            split_k=split)
    finally:
        ctx.config("mfma_tile", 0)
    got = dC.read(np.int32, b * d).reshape(b, d)
    np.testing.assert_array_equal(got, oracle_lib.popcount_matrix(A, B))
    for x in (dA, dB, dC):
        x.free()


@pytest.mark.parametrize("path", ["mfma", "fp4", "packed"])
def test_popcount_matrix_all_ones_largest_split(ctx, path):
    """Rows of ones over 2^24 + 2,048 + 1 bits: the first word split holds the
    largest number of bits a split may hold (2^24), so the matrix-core
    accumulators reach their extreme (int8: 64 x 2^24 = 2^30; fp4: 2 x 2^24 =
    2^25 in fp32, still exact); every count must be the row length, and a row
    of zeros must give 0."""
    n_bits = (1 << 24) + 2048 + 1
    pitch = tf.pitch_words(n_bits)
    b, d = 3, 2
    ones = rand_rows(np.random.default_rng(0), 1, n_bits, 1.0)[0]
    A = np.zeros((b, pitch), np.uint64)
    B = np.zeros((d, pitch), np.uint64)
    w = (n_bits + 63) // 64
    A[0, :w] = ones
    A[2, :w] = ones
    B[:, :w] = ones
    dA, dB, dC = ctx.buffer(A.nbytes), ctx.buffer(B.nbytes), ctx.buffer(b * d * 4)
    dA.write(A)
    dB.write(B)
    ctx.popcount_matrix(dA, dB, dC, b, d, pitch, path=path, split_k=1)
    got = dC.read(np.int32, b * d).reshape(b, d)
    np.testing.assert_array_equal(got, np.array([[n_bits] * d, [0] * d, [n_bits] * d], np.int32))
    for x in (dA, dB, dC):
        x.free()


@pytest.mark.parametrize("graphs", [1, 0])
@pytest.mark.parametrize("kernel", [1, 2, 3, 4, 5, "4-int8", "5-int8"])
@pytest.mark.parametrize("n_bits,n_items", [(1100, 30), (3200, 75), (400_003, 90), (2_000_000, 38)])
def test_wave_kernels_forced(ctx, kernel, graphs, n_bits, n_items):
    """Every batched path (1 = L2-resident kernel, 2 = LDS-staged, 3 =
    LDS-staged with wide staging, 4 = matrix cores: masks, fp4 matrix count of
    all (row, survivor) pairs, gather; 5 = the fused matrix-core kernel; both
    matrix-core forms also on the int8 matrix cores) on
    the same waves,
    replayed from graphs or issued directly: uneven prefixes, uneven and empty
    tails, full-column rows."""
    int8 = isinstance(kernel, str)   # the matrix-core paths on the int8 matrix cores
    kernel = int(str(kernel)[0])
    rng = np.random.default_rng(kernel * 100 + n_items + 7 * int8)
    rows = rand_rows(rng, n_items, n_bits, 0.65)
    n_cols = n_items - 1
    ctx.load_bitmaps(rows, n_bits, n_cols)
    ctx.config("wave_kernel", kernel)
    ctx.config("graphs", graphs)
    ctx.config("mfma_format", 1 if int8 else 0)
    try:
        prefixes, tails = [], []
        for i in range(120):
            prefixes.append(rng.choice(n_items, size=int(rng.integers(1, 9)), replace=False))
            t = int(rng.integers(0, 40)) if i % 7 else 0
            tails.append(np.sort(rng.choice(n_cols, size=min(t, n_cols), replace=False)))
        got = ctx.count_nodes(prefixes, tails)
        for p, t, g in zip(prefixes, tails, got):
            np.testing.assert_array_equal(g, oracle_lib.node_counts(rows, p, n_cols)[t])
        full = ctx.count_nodes(prefixes[:60])
        for p, g in zip(prefixes[:60], full):
            np.testing.assert_array_equal(g, oracle_lib.node_counts(rows, p, n_cols))
    finally:
        ctx.config("wave_kernel", 0)
        ctx.config("graphs", 0)
        ctx.config("mfma_format", 0)


@pytest.mark.parametrize("n_rows", [250, 129])
def test_wave_matrix_cores_large_waves(ctx, n_rows):
    """Matrix-core waves (fp4 count) of more than 128 rows against 130
    survivors (two column tiles) over a row length off every multiple of 2,048
    bits, with uneven prefixes and tails and full-column rows, against the
    oracle."""
    rng = np.random.default_rng(n_rows)
    n_bits, n_items = 3 * 2048 * 7 + 77, 131
    rows = rand_rows(rng, n_items, n_bits, 0.7)
    n_cols = n_items - 1
    ctx.load_bitmaps(rows, n_bits, n_cols)
    ctx.config("wave_kernel", 4)
    try:
        prefixes = [rng.choice(n_items, size=int(rng.integers(1, 7)), replace=False) for _ in range(n_rows)]
        tails = [np.sort(rng.choice(n_cols, size=int(rng.integers(0, 60)), replace=False)) for _ in range(n_rows)]
        got = ctx.count_nodes(prefixes, tails)
        for p, t, g in zip(prefixes, tails, got):
            np.testing.assert_array_equal(g, oracle_lib.node_counts(rows, p, n_cols)[t])
        full = ctx.count_nodes(prefixes)
        for p, g in zip(prefixes, full):
            np.testing.assert_array_equal(g, oracle_lib.node_counts(rows, p, n_cols))
    finally:
        ctx.config("wave_kernel", 0)


def test_wave_graph_replays(ctx):
    """Graph replays with changing wave sizes, interleaved with eager work and
    host synchronisation, each checked against the oracle."""
    rng = np.random.default_rng(77)
    n_bits, n_items = 50_001, 40
    rows = rand_rows(rng, n_items, n_bits, 0.6)
    ctx.load_bitmaps(rows, n_bits, n_items)
    ctx.config("graphs", 1)
    before = ctx.wave_stats()
    for rep, n in enumerate([50, 3, 200, 1, 120, 7, 60]):
        prefixes = [rng.choice(n_items, size=int(rng.integers(1, 6)), replace=False) for _ in range(n)]
        tails = [np.sort(rng.choice(n_items, size=int(rng.integers(0, 25)), replace=False)) for _ in range(n)]
        got = ctx.count_nodes(prefixes, tails)
        for p, t, g in zip(prefixes, tails, got):
            np.testing.assert_array_equal(g, oracle_lib.node_counts(rows, p, n_items)[t], err_msg="replay %d" % rep)
        # eager work between replays
        np.testing.assert_array_equal(ctx.count_node(prefixes[0]), oracle_lib.node_counts(rows, prefixes[0], n_items))
    after = ctx.wave_stats()
    ctx.config("graphs", 0)
    assert after["graph_launches"] - before["graph_launches"] >= 7
    assert after["graphs"] - before["graphs"] <= 8   # instantiated once per (slot, kernel)


@pytest.mark.parametrize("path", ["mfma", "packed", "fp4"])
def test_popcount_matrix_tuned_split(ctx, path):
    """split_k = 0: the tuner times the candidates on first use of the shape,
    keeps its choice, and the result stays exact."""
    rng = np.random.default_rng(99)
    b, d, n_bits = 300, 140, 2048 * 40 + 5
    pitch = tf.pitch_words(n_bits)
    w = (n_bits + 63) // 64
    A = np.zeros((b, pitch), np.uint64)
    B = np.zeros((d, pitch), np.uint64)
    A[:, :w] = rand_rows(rng, b, n_bits, 0.4)
    B[:, :w] = rand_rows(rng, d, n_bits, 0.4)
    dA, dB, dC = ctx.buffer(A.nbytes), ctx.buffer(B.nbytes), ctx.buffer(b * d * 4)
    dA.write(A)
    dB.write(B)
    exp = oracle_lib.popcount_matrix(A, B)
    for _ in range(2):   # first call tunes, second reuses the choice
        ctx.popcount_matrix(dA, dB, dC, b, d, pitch, path=path, split_k=0)
        np.testing.assert_array_equal(dC.read(np.int32, b * d).reshape(b, d), exp)
    tile, split = ctx.popcount_matrix_plan(dA, dB, dC, b, d, pitch, path=path)
    assert split in (1, 2, 4, 8, 16, 32, 64, 128, 256, 512) and split * 64 <= pitch or split == 1
    assert tile in {"packed": (64,), "fp4": (256,)}.get(path, (128, 256))
    for x in (dA, dB, dC):
        x.free()


@pytest.mark.parametrize("path", ["packed", "mfma"])
def test_count_nodes_path_argument(ctx, path):
    """The path argument of a wave: packed kernels only, or the matrix cores;
    both exact, on rows off multiples of 64 and 2048 bits."""
    rng = np.random.default_rng(5 if path == "packed" else 6)
    n_bits, n_items = 2048 * 11 + 77, 140
    rows = rand_rows(rng, n_items, n_bits, 0.6)
    ctx.load_bitmaps(rows, n_bits, n_items - 1)
    before = ctx.wave_stats()
    prefixes = [rng.choice(n_items, size=int(rng.integers(1, 6)), replace=False) for _ in range(90)]
    tails = [np.sort(rng.choice(n_items - 1, size=int(rng.integers(0, 60)), replace=False)) for _ in range(90)]
    got = ctx.count_nodes(prefixes, tails, path=path)
    for p, t, g in zip(prefixes, tails, got):
        np.testing.assert_array_equal(g, oracle_lib.node_counts(rows, p, n_items - 1)[t])
    after = ctx.wave_stats()
    mat = lambda st: st["mfma_launches"] + st["fused_launches"]   # noqa: E731
    if path == "mfma":
        assert after["fused_launches"] == before["fused_launches"] + 1   # the fused kernel takes it
        assert mat(after) == mat(before) + 1
    else:
        assert mat(after) == mat(before)


@pytest.mark.parametrize("n_bits,n_items,n_cols,n_rows,kmax", [
    (2048 * 3 + 5, 60, 59, 700, 10),        # 64-column tiles, 256-row tiles, three row tiles
    (300_001, 300, 290, 300, 6),            # 128-column tiles, three of them, 128-row tiles
    (1_500_007, 40, 39, 150, 25),           # long rows: the word axis split, deep prefixes
    (5_000, 260, 200, 64, 3)])              # one 64-row tile
def test_fused_matrix_core_wave_shapes(ctx, n_bits, n_items, n_cols, n_rows, kmax):
    """The fused matrix-core kernel (wave_kernel 5) over several row and
    column tiles, split word ranges and deep prefixes; tails and full rows."""
    rng = np.random.default_rng(n_rows + n_items)
    rows = rand_rows(rng, n_items, n_bits, 0.7)
    ctx.load_bitmaps(rows, n_bits, n_cols)
    ctx.config("wave_kernel", 5)
    try:
        before = ctx.wave_stats()["fused_launches"]
        for rep in range(2):
            pool = rng.choice(n_items, size=min(n_items, 250), replace=False)   # <= 256 distinct prefix items
            prefixes = [rng.choice(pool, size=int(rng.integers(1, kmax + 1)), replace=False) for _ in range(n_rows)]
            tails = [np.sort(rng.choice(n_cols, size=int(rng.integers(0, n_cols + 1)), replace=False))
                     for _ in range(n_rows)]
            print("rep %d: %d rows, %d distinct prefix items" % (rep, n_rows, len(set(np.concatenate(prefixes)))),
                  flush=True)
            got = ctx.count_nodes(prefixes, tails)
            for i, (p, t, g) in enumerate(zip(prefixes, tails, got)):
                np.testing.assert_array_equal(g, oracle_lib.node_counts(rows, p, n_cols)[t], err_msg="row %d" % i)
        full = ctx.count_nodes(prefixes[:100])
        for p, g in zip(prefixes[:100], full):
            np.testing.assert_array_equal(g, oracle_lib.node_counts(rows, p, n_cols))
        assert ctx.wave_stats()["fused_launches"] == before + 3
    finally:
        ctx.config("wave_kernel", 0)


def test_fused_matrix_core_limits(ctx):
    """A wave with 257 distinct prefix items is outside the fused kernel:
    forcing it is an error, and the matrix-core path argument then takes the
    three-launch form, exactly."""
    rng = np.random.default_rng(3)
    n_bits, n_items, n_cols = 9_000, 300, 280
    rows = rand_rows(rng, n_items, n_bits, 0.8)
    ctx.load_bitmaps(rows, n_bits, n_cols)
    prefixes = [[i] for i in range(257)]
    tails = [list(range(0, n_cols, 7))] * 257
    ctx.config("wave_kernel", 5)
    try:
        with pytest.raises(tf.TfimError, match="256 distinct prefix items"):
            ctx.count_nodes(prefixes, tails)
    finally:
        ctx.config("wave_kernel", 0)
    before = ctx.wave_stats()
    got = ctx.count_nodes(prefixes, tails, path="mfma")
    for p, t, g in zip(prefixes, tails, got):
        np.testing.assert_array_equal(g, oracle_lib.node_counts(rows, p, n_cols)[t])
    after = ctx.wave_stats()
    assert after["mfma_launches"] == before["mfma_launches"] + 1
    assert after["fused_launches"] == before["fused_launches"]
