"""GPU parity of the isolated support count at its full size: 65,536 prefix
bitmaps x 512 item bitmaps over 2^22 transactions, i.i.d. Bernoulli(0.3) bits,
seed 1 (the microbench shape the headline throughput is quoted on).

A is 32 GiB, so its rows past 8,192 sit beyond 2^32 bytes, and the count
kernel's grid is 1,024 x 8 workgroups at split-k 8.  The reference gates its
microbench the same way: every GPU kernel against the CPU reference on the
gate shape (code/microbench/bmma_bench.cu:326-358,431-434).

Checked rows (all 512 columns each): the first and the last 256-row tile,
the rows on both sides of every 4 GiB boundary of A (rows 8,192k - 1 and
8,192k), and 64 random rows.  The expected counts come from A and B rows
regenerated on the CPU from the seed by the oracle's generator, so the
device generator is pinned too.
"""
import numpy as np
import pytest
from concurrent.futures import ThreadPoolExecutor

import oracle_lib
import tensorfim_amd as tf

pytestmark = pytest.mark.gpu

ROWS, ITEMS, N_BITS, SEED = 65536, 512, 1 << 22, 1
THR32 = int(np.floor(0.3 * 2.0 ** 32))


def sample_rows():
    rows = set(range(256)) | set(range(ROWS - 256, ROWS))
    row_bytes = tf.pitch_words(N_BITS) * 8
    per4g = (1 << 32) // row_bytes
    for k in range(1, ROWS // per4g):
        rows |= {k * per4g - 1, k * per4g}
    rng = np.random.default_rng(12)
    rows |= set(int(r) for r in rng.choice(ROWS, size=64, replace=False))
    return np.array(sorted(rows), dtype=np.int64)


def cpu_rows(matrix, rows, pitch):
    with ThreadPoolExecutor(16) as ex:
        return np.stack(list(ex.map(lambda r: oracle_lib.bernoulli_row(SEED, matrix, int(r), N_BITS, pitch, THR32),
                                    rows)))


@pytest.fixture(scope="module")
def full():
    pitch = tf.pitch_words(N_BITS)
    rows = sample_rows()
    print("isolated full size: %d x %d x %d bits, %d sampled rows" % (ROWS, ITEMS, N_BITS, len(rows)), flush=True)
    ctx = tf.Context(0)
    A = ctx.buffer(ROWS * pitch * 8)
    B = ctx.buffer(ITEMS * pitch * 8)
    C = ctx.buffer(ROWS * ITEMS * 4)
    ctx.fill_bernoulli(A, ROWS, N_BITS, pitch, seed=SEED, matrix=0, p=0.3)
    ctx.fill_bernoulli(B, ITEMS, N_BITS, pitch, seed=SEED, matrix=1, p=0.3)
    ctx.sync()
    A_cpu = cpu_rows(0, rows, pitch)
    B_cpu = cpu_rows(1, np.arange(ITEMS), pitch)
    expected = oracle_lib.popcount_matrix(A_cpu, B_cpu, threads=16)
    yield ctx, A, B, C, pitch, rows, A_cpu, B_cpu, expected
    for x in (A, B, C):
        x.free()
    ctx.close()


def read_rows(buf, rows, width, dtype, pitch_bytes):
    return np.stack([buf.read(dtype, width, offset=int(r) * pitch_bytes) for r in rows])


def test_device_generator_matches_oracle(full):
    ctx, A, B, C, pitch, rows, A_cpu, B_cpu, _ = full
    np.testing.assert_array_equal(read_rows(A, rows, pitch, np.uint64, pitch * 8), A_cpu)
    np.testing.assert_array_equal(B.read(np.uint64, ITEMS * pitch).reshape(ITEMS, pitch), B_cpu)
    assert abs(float(np.mean([np.unpackbits(r.view(np.uint8)).mean() for r in A_cpu[:8]])) * 2 ** 22 / N_BITS
               - 0.3) < 0.01


@pytest.mark.parametrize("path,split", [("mfma", 8), ("fp4", 8), ("packed", 1), ("mfma", 1), ("fp4", 1),
                                        ("packed", 8)])
def test_full_size_counts_match_oracle(full, path, split):
    ctx, A, B, C, pitch, rows, _, _, expected = full
    print("path %s split_k %d" % (path, split), flush=True)
    ctx.popcount_matrix(A, B, C, ROWS, ITEMS, pitch, path=path, split_k=split)
    ctx.sync()
    got = read_rows(C, rows, ITEMS, np.int32, ITEMS * 4)
    bad = np.argwhere(got != expected)
    assert len(bad) == 0, "path %s split %d: %d mismatches, first at row %d col %d" % (
        path, split, len(bad), rows[bad[0][0]], bad[0][1])


def test_full_size_tuned_plan_is_exact(full):
    """The library's own choice of tile and split (split_k = 0) on this
    shape."""
    ctx, A, B, C, pitch, rows, _, _, expected = full
    tile, split = ctx.popcount_matrix_plan(A, B, C, ROWS, ITEMS, pitch, path="mfma")
    print("tuned: tile %d split %d" % (tile, split), flush=True)
    ctx.popcount_matrix(A, B, C, ROWS, ITEMS, pitch, path="mfma", split_k=0)
    ctx.sync()
    np.testing.assert_array_equal(read_rows(C, rows, ITEMS, np.int32, ITEMS * 4), expected)
