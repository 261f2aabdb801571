"""CPU tests: the oracle pinned against the reference's own golden data and
known answers, plus closed-form checks a plausible mistake would fail."""
import itertools
import os
import subprocess

import numpy as np
import pytest

import oracle_lib
import tensorfim_amd as tf
from conftest import GOLD_PAIRS, gold_path


@pytest.mark.parametrize("ds,thr", GOLD_PAIRS)
def test_oracle_reproduces_gold(fimi_dir, tmp_path, ds, thr):
    out = tmp_path / "o.res"
    oracle_lib.mine(fimi_dir / ds, thr, threads=8, out=out)
    assert tf.load_results(str(out)) == tf.load_results(gold_path(ds, thr))
    h, g = tf.load_header(str(out)), tf.load_header(gold_path(ds, thr))
    for k in ["transNum", "dimension", "itemIDmin", "itemIDmax", "transLengthMin", "transLengthMax",
              "transLengthMean", "dimensionReduced", "transNumReduced", "transLengthReducedMin",
              "transLengthReducedMax", "transLengthReducedMean", "MFIsNumber"]:
        assert h[k] == g[k], k


@pytest.mark.slow
def test_batch_prototype_known_answer():
    """batch_proto.cu generator + level-wise counting: F1=248 and per-level
    counts 585, 587, 350, 117, 17, 0 (1,904 total; run_dispatch.txt:3-16)."""
    import subprocess
    oracle_lib.build()
    r = subprocess.run([oracle_lib.CLI, "proto", "--threads", "8"], capture_output=True, text=True, check=True)
    lv = {ln.split()[0]: int(ln.split()[1]) for ln in r.stdout.splitlines() if ln.startswith("L")}
    assert [lv["L%d" % i] for i in range(1, 8)] == [248, 585, 587, 350, 117, 17, 0]
    assert "total 1904" in r.stdout


def _brute_mfis(rows, support):
    """Reference semantics by brute force over every itemset (tiny inputs)."""
    trans_num = len(rows)
    thr = trans_num * support
    items = sorted({x for r in rows for x in r})
    freq = {x: sum(r.count(x) for r in rows) for x in items}
    surv = [x for x in items if x >= 1 and not freq[x] < thr]
    # roots: strict relative test
    roots = {x for x in surv if freq[x] / trans_num > support}
    sets = [set(r) for r in rows]
    sup = lambda s: sum(1 for r in sets if s <= r)
    out = {}
    for k in range(1, len(surv) + 1):
        for c in itertools.combinations(surv, k):
            s = set(c)
            v = sup(s)
            if v < thr or max(c) not in roots:
                continue
            if any(sup(s | {e}) >= thr for e in surv if e not in s):
                continue
            out[frozenset(c)] = freq[c[0]] if k == 1 else v
    return out


@pytest.mark.parametrize("seed", range(12))
def test_oracle_matches_brute_force(tmp_path, seed):
    rng = np.random.default_rng(seed)
    n_items = int(rng.integers(4, 10))
    rows = [sorted(set(rng.choice(np.arange(1, n_items + 1), size=int(rng.integers(1, n_items + 1)),
                                  replace=True).tolist())) for _ in range(int(rng.integers(5, 40)))]
    support = [0.2, 0.3, 0.5, 0.25][seed % 4]
    if seed % 3 == 0:
        rows.append([])        # empty line: counts in transNum (quirk A.1)
    path = tmp_path / "db.txt"
    path.write_text("\n".join(" ".join(map(str, r)) for r in rows) + "\n")
    out = tmp_path / "o.res"
    oracle_lib.mine(path, support, threads=2, out=out)
    assert tf.load_results(str(out)) == _brute_mfis(rows, support)


def test_exact_threshold_items_are_never_roots(tmp_path):
    """Quirk A.2: an item at exactly transNum*support survives but is not a
    root, so the MFI {2} (max id 2, support exactly 2 of 4) is not reported,
    while {1} is reported only if it has no frequent extension."""
    path = tmp_path / "q.txt"
    path.write_text("1 2\n1\n1 2\n1\n")
    out = tmp_path / "o.res"
    oracle_lib.mine(path, 0.5, threads=1, out=out)
    # {1,2}: support 2 >= 2, max id 2 has freq 2 == thr -> not a root -> dropped
    assert tf.load_results(str(out)) == {}


def test_popcount_matrix_closed_form():
    rng = np.random.default_rng(3)
    A = rng.integers(0, 2 ** 63, size=(7, 33), dtype=np.uint64)
    B = rng.integers(0, 2 ** 63, size=(5, 33), dtype=np.uint64)
    C = oracle_lib.popcount_matrix(A, B, threads=2)
    for i in range(7):
        for j in range(5):
            assert C[i, j] == sum(bin(int(a) & int(b)).count("1") for a, b in zip(A[i], B[j]))
    # all-ones row gives item supports; symmetric when A == B
    ones = np.full((1, 33), np.uint64(2 ** 64 - 1))
    assert (oracle_lib.popcount_matrix(ones, B)[0] == [sum(bin(int(x)).count("1") for x in b) for b in B]).all()
    S = oracle_lib.popcount_matrix(A, A)
    assert (S == S.T).all()


def test_bernoulli_rng_density_and_tail():
    thr32 = int(0.3 * 2 ** 32)
    row = oracle_lib.bernoulli_row(1, 0, 5, 100_003, 1563, thr32)
    bits = np.unpackbits(row.view(np.uint8), bitorder="little")
    assert bits[100_003:].sum() == 0
    assert abs(bits[:100_003].mean() - 0.3) < 0.01
    other = oracle_lib.bernoulli_row(1, 1, 5, 100_003, 1563, thr32)
    assert (row != other).any()


@pytest.mark.parametrize("name", ["dense-small", "gene"])
def test_synthetic_fixture_reproduces(tmp_path, name):
    """The committed golden digests of the small synthetic workloads are what
    the generator and the oracle give now (the larger ones take minutes and
    are checked on the GPU against the engines)."""
    import json
    from conftest import mfi_digest
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "synthetic", name + ".json")) as f:
        fx = json.load(f)
    data = tmp_path / (name + ".txt")
    subprocess.run([tf.binary("tfim_gen"), fx["workload"], str(data)], check=True)
    res = tmp_path / "o.res"
    oracle_lib.mine(data, fx["support"], threads=4, out=res)
    got = {s: f * fx["replicate"] for s, f in tf.load_results(str(res)).items()}
    assert len(got) == fx["n_mfi"]
    assert mfi_digest(got) == fx["digest"]
