"""GPU parity of complement rows (tfim_count_nodes_ex): a row counted against
every survivor outside an exclusion bitmap, with the threshold step done on
the device, returns (support of the prefix, number of frequent non-excluded
survivors other than its first prefix item).  Expected values come from the
oracle's counts of every survivor (the reference's maximality test,
kernel_bmma.cu:857-877, counts all survivors and thresholds them on the
host).  Waves mix explicit and complement rows; sizes cover several word
chunks, survivor counts off multiples of 64, prefix items excluded or not,
and thresholds of 0, typical and above every count.
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


@pytest.fixture(scope="module")
def ctx():
    c = tf.Context(0)
    yield c
    c.close()


def expected(rows, n_cols, prefix, excl, min_count):
    cnt = oracle_lib.node_counts(rows, prefix, n_cols)
    x0 = prefix[0]
    live = [e for e in range(n_cols) if e != x0 and e not in excl]
    return np.array([cnt[x0], sum(1 for e in live if cnt[e] >= min_count)], dtype=np.int32)


def make_wave(rng, n_items, n_cols, n_rows, p_compl):
    prefixes, tails, excl = [], [], []
    for _ in range(n_rows):
        k = int(rng.integers(1, 5))
        if rng.random() < p_compl:
            # complement row: first prefix item is a survivor
            first = int(rng.integers(0, n_cols))
            rest = [int(x) for x in rng.choice(n_items, size=k - 1, replace=False) if x != first]
            prefix = [first] + rest
            ex = set(int(x) for x in rng.choice(n_cols, size=int(rng.integers(0, n_cols)), replace=False))
            if rng.random() < 0.7:
                ex |= {p for p in prefix if p < n_cols}
            prefixes.append(prefix)
            tails.append(None)
            excl.append(ex)
        else:
            prefixes.append([int(x) for x in rng.choice(n_items, size=k, replace=False)])
            nt = int(rng.integers(0, min(n_cols, 90)))
            tails.append([int(x) for x in rng.choice(n_cols, size=nt, replace=False)])
            excl.append(None)
    return prefixes, tails, excl


@pytest.mark.parametrize("kernel", [1, 4, 5])
@pytest.mark.parametrize("n_bits,n_items,n_cols", [(70, 9, 8), (3200, 131, 130), (2048 * 5 + 17, 200, 199),
                                                   (300_017, 150, 128), (1_000_003, 70, 65)])
def test_count_nodes_ex_matches_oracle(ctx, n_bits, n_items, n_cols, kernel):
    """Kernel 1: L2 kernel + complement reduction; kernel 4: matrix cores
    (masks, int8-MFMA count of every pair, gather and reduction); kernel 5:
    the fused matrix-core kernel (complement rows thresholded by the last
    workgroup of each row tile)."""
    rng = np.random.default_rng(n_bits + n_items)
    rows = rand_rows(rng, n_items, n_bits, 0.6)
    ctx.load_bitmaps(rows, n_bits, n_cols)
    ctx.config("wave_kernel", kernel)
    try:
        _check_ex_waves(ctx, rng, rows, n_bits, n_items, n_cols)
    finally:
        ctx.config("wave_kernel", 0)


def _check_ex_waves(ctx, rng, rows, n_bits, n_items, n_cols):
    for p_compl in (1.0, 0.5):
        prefixes, tails, excl = make_wave(rng, n_items, n_cols, 40, p_compl)
        sups = [int(oracle_lib.node_counts(rows, p, n_cols)[p[0]]) if t is None else 0
                for p, t in zip(prefixes, tails)]
        typical = int(np.median([s for s in sups if s] or [1])) // 2 + 1
        for min_count in (0, typical, n_bits + 1):
            print("n_bits=%d n_cols=%d p_compl=%.1f min_count=%d" % (n_bits, n_cols, p_compl, min_count), flush=True)
            got = ctx.count_nodes_ex(prefixes, tails, excl, min_count)
            for i, (p, t, ex) in enumerate(zip(prefixes, tails, excl)):
                if t is None:
                    np.testing.assert_array_equal(got[i], expected(rows, n_cols, p, ex, min_count),
                                                  err_msg="row %d prefix %s" % (i, p))
                else:
                    np.testing.assert_array_equal(got[i], oracle_lib.node_counts(rows, p, n_cols)[t],
                                                  err_msg="row %d prefix %s" % (i, p))


@pytest.mark.parametrize("graphs", [1, 0])
def test_complement_waves_replayed(ctx, graphs):
    """Waves of changing size alternate with and without complement rows (two
    graph variants per slot), with eager checks between them, on the
    automatic choice between the L2 kernel and the matrix cores."""
    rng = np.random.default_rng(11 + graphs)
    n_bits, n_items, n_cols = 40_000, 300, 300
    rows = rand_rows(rng, n_items, n_bits, 0.7)
    ctx.load_bitmaps(rows, n_bits, n_cols)
    ctx.config("graphs", graphs)
    try:
        for it in range(8):
            n_rows = [50, 7, 120, 1, 64, 33, 2, 90][it]
            prefixes, tails, excl = make_wave(rng, n_items, n_cols, n_rows, 0.0 if it % 3 == 2 else 0.6)
            min_count = n_bits // 4 + it
            if all(t is not None for t in tails):
                got = ctx.count_nodes(prefixes, tails)
            else:
                got = ctx.count_nodes_ex(prefixes, tails, excl, min_count)
            for i, (p, t, ex) in enumerate(zip(prefixes, tails, excl)):
                exp = expected(rows, n_cols, p, ex, min_count) if t is None else \
                    oracle_lib.node_counts(rows, p, n_cols)[t]
                np.testing.assert_array_equal(got[i], exp, err_msg="wave %d row %d" % (it, i))
        st = ctx.wave_stats()
        print("paths: l2 %d mfma %d" % (st["l2_launches"], st["mfma_launches"]), flush=True)
    finally:
        ctx.config("graphs", 0)


def test_complement_rows_rejected_when_malformed(ctx):
    rng = np.random.default_rng(5)
    rows = rand_rows(rng, 12, 500, 0.5)
    ctx.load_bitmaps(rows, 500, 10)
    with pytest.raises(tf.TfimError):
        ctx.count_nodes_ex([[11, 2]], [None], [set()], 3)   # first item is not a survivor
    ok = ctx.count_nodes_ex([[3, 11]], [None], [{3}], 3)
    np.testing.assert_array_equal(ok[0], expected(rows, 10, [3, 11], {3}, 3))


@pytest.mark.parametrize("kernel", [4, 5])
def test_matrix_core_split_tuner_arms(ctx, kernel):
    """Matrix-core waves on long rows (split over the words) run every arm of
    the split tuner (1/2, 1 and 2 times the occupancy rule's split, three
    samples each, then re-tries): every wave stays exact, in the three-launch
    form (4) and the fused kernel (5)."""
    rng = np.random.default_rng(23)
    n_bits, n_items, n_cols = 1_500_007, 60, 57
    rows = rand_rows(rng, n_items, n_bits, 0.55)
    ctx.load_bitmaps(rows, n_bits, n_cols)
    ctx.config("wave_kernel", kernel)
    try:
        for it in range(14):
            prefixes, tails, excl = make_wave(rng, n_items, n_cols, int(rng.integers(1, 70)), 0.4)
            min_count = int(n_bits * 0.2) + it
            got = ctx.count_nodes_ex(prefixes, tails, excl, min_count)
            for i, (p, t, ex) in enumerate(zip(prefixes, tails, excl)):
                exp = expected(rows, n_cols, p, ex, min_count) if t is None else \
                    oracle_lib.node_counts(rows, p, n_cols)[t]
                np.testing.assert_array_equal(got[i], exp, err_msg="wave %d row %d" % (it, i))
    finally:
        ctx.config("wave_kernel", 0)
