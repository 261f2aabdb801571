"""GPU parity of the mask-row batched path (tfim_count_masks) against the CPU
oracle (bit-exact).

Rows are (prefix ids, column ids) given as 64-bit masks, for databases of at
most 64 dense ids; the counts of a row come back in ascending column order.
Shapes: transaction counts off multiples of 64 and 2048, 9 to 64 ids, a
prefix-only last id (item 0, kernel_bmma.cu:402-403), rows of 1 to 63
columns (full groups of eight, and every remainder 1 .. 7 that the kernel
pads), empty column masks, every staging and tile arm, graphs on and off,
and waves beyond the kernel's 8,192 outputs (counted through id lists).
"""
import numpy as np
import pytest

import oracle_lib
import tensorfim_amd as tf

pytestmark = pytest.mark.gpu


def rand_rows(rng, n_items, n_bits, density):
    words = (n_bits + 63) // 64
    bits = rng.random((n_items, words * 64)) < density
    bits[:, n_bits:] = False
    return np.packbits(bits, axis=1, bitorder="little").view(np.uint64).reshape(n_items, words)


def rand_wave(rng, n_items, n_cols, n_rows, kmax=6):
    prefixes, cols = [], []
    for r in range(n_rows):
        k = int(rng.integers(1, kmax + 1))
        prefixes.append(sorted(rng.choice(n_items, size=min(k, n_items), replace=False).tolist()))
        nc = int(rng.integers(0, n_cols + 1)) if r % 5 else (r // 5) % (n_cols + 1)
        cols.append(sorted(rng.choice(n_cols, size=nc, replace=False).tolist()))
    return prefixes, cols


def check(ctx, rows, n_cols, prefixes, cols, tag):
    got = ctx.count_masks(prefixes, cols)
    for i, (p, c) in enumerate(zip(prefixes, cols)):
        exp = oracle_lib.node_counts(rows, p, n_cols)[c] if c else np.zeros(0, np.int32)
        np.testing.assert_array_equal(got[i], exp, err_msg="%s row %d prefix %s cols %s" % (tag, i, p, c))


@pytest.fixture(scope="module")
def ctx():
    c = tf.Context(0)
    print("device:", c.device_info(), flush=True)
    yield c
    c.close()


@pytest.mark.parametrize("n_bits,n_items,n_cols", [(70, 9, 8), (3200, 39, 39), (2048 * 5 + 17, 64, 63),
                                                   (1_000_003, 40, 39)])
@pytest.mark.parametrize("arm", ["auto", "narrow16", "narrow32", "narrow64", "wide16", "wide32", "wide64"])
def test_masks_match_oracle(ctx, n_bits, n_items, n_cols, arm):
    print("masks", n_bits, n_items, n_cols, arm, flush=True)
    rng = np.random.default_rng(n_bits + n_items)
    rows = rand_rows(rng, n_items, n_bits, 0.6)
    ctx.load_bitmaps(rows, n_bits, n_cols)
    if arm != "auto":
        ctx.config("wave_kernel", 2 if arm.startswith("narrow") else 3)
        ctx.config("tile_cols", int(arm[-2:]))
    try:
        for rep in range(3):
            prefixes, cols = rand_wave(rng, n_items, n_cols, 150)
            check(ctx, rows, n_cols, prefixes, cols, "%s rep %d" % (arm, rep))
    finally:
        ctx.config("wave_kernel", 0)
        ctx.config("tile_cols", 0)


@pytest.mark.parametrize("graphs", [0, 1])
def test_masks_graphs_and_large_waves(ctx, graphs):
    """Graph replays and direct launches agree with the oracle; a wave of more
    than 8,192 outputs (counted through id lists) too."""
    print("masks graphs", graphs, flush=True)
    rng = np.random.default_rng(99 + graphs)
    n_bits, n_items, n_cols = 12_345, 40, 40
    rows = rand_rows(rng, n_items, n_bits, 0.7)
    ctx.load_bitmaps(rows, n_bits, n_cols)
    ctx.config("graphs", graphs)
    try:
        for n_rows in (5, 300, 40, 900):
            prefixes, cols = rand_wave(rng, n_items, n_cols, n_rows, kmax=12)
            check(ctx, rows, n_cols, prefixes, cols, "graphs %d rows %d" % (graphs, n_rows))
    finally:
        ctx.config("graphs", 0)


def test_masks_refuse_bad_input(ctx):
    rng = np.random.default_rng(5)
    rows = rand_rows(rng, 70, 1000, 0.5)
    ctx.load_bitmaps(rows, 1000, 69)
    with pytest.raises(RuntimeError, match="64"):
        ctx.count_masks([[1]], [[2]])          # more than 64 ids loaded
    rows = rand_rows(rng, 30, 1000, 0.5)
    ctx.load_bitmaps(rows, 1000, 20)
    with pytest.raises(RuntimeError, match="column mask"):
        ctx.count_masks([[1]], [[25]])         # column beyond n_cols
    with pytest.raises(RuntimeError, match="prefix mask"):
        ctx.count_masks([[]], [[2]])           # empty prefix
    np.testing.assert_array_equal(ctx.count_masks([[3, 29]], [[0, 19]])[0],
                                  oracle_lib.node_counts(rows, [3, 29], 20)[[0, 19]])
