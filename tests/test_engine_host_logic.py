"""CPU: the engine's search logic (plain, wave and MAFIA engines, policies,
quirks) against the oracle, with the counting library replaced by a CPU
stand-in (tests/engine_mock/mock_tfim.cpp) so it runs without a GPU.  The
GPU tests run the same engines on the real library."""
import os
import random
import subprocess

import pytest

import oracle_lib
import tensorfim_amd as tf
from conftest import GOLD_PAIRS, gold_path

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ENG = os.path.join(ROOT, "tensorfim_amd", "csrc", "engine")


@pytest.fixture(scope="module")
def mock_engine(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("mock") / "tfim_mine_mock")
    srcs = [os.path.join(ENG, f) for f in ("dataset.cpp", "results.cpp", "search.cpp", "policy.cpp", "main_mine.cpp")]
    subprocess.run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), "-I" + ENG] + srcs +
                   [os.path.join(ROOT, "tests", "engine_mock", "mock_tfim.cpp"), "-o", out, "-lpthread"], check=True)
    return out


ENGINES = [["--engine", "mafia"], ["--engine", "mafia", "--gate", "always"], ["--engine", "wave"],
           ["--engine", "plain", "--workers", "3"], ["--engine", "mafia", "--compl", "on"],
           ["--engine", "mafia", "--gate", "always", "--compl", "on"], ["--engine", "mafia", "--spec", "off"],
           ["--engine", "mafia", "--compl", "on", "--spec", "off"], ["--engine", "mafia", "--rows", "lists"],
           ["--engine", "mafia", "--learned"], ["--engine", "mafia", "--learned", "--compl", "on"],
           ["--engine", "mafia", "--learned", "--rank"]]


@pytest.mark.parametrize("ds,thr", GOLD_PAIRS[:5])
def test_engines_reproduce_golds(mock_engine, fimi_dir, tmp_path, ds, thr):
    for eng in ENGINES[:3] + ENGINES[4:]:
        out = tmp_path / "m.res"
        subprocess.run([mock_engine, str(fimi_dir / ds), repr(thr), "--out", str(out)] + eng, check=True,
                       capture_output=True)
        assert tf.load_results(str(out)) == tf.load_results(gold_path(ds, thr)), eng


def test_engines_match_oracle_on_random_databases(mock_engine, tmp_path):
    """Random small databases with frequent item 0 (its own enumeration rule),
    repeated ids, empty lines and exact-threshold items."""
    rnd = random.Random(7)
    for it in range(80):
        n_items = rnd.randint(3, 11)
        rows = []
        for _ in range(rnd.randint(4, 50)):
            r = rnd.sample(range(0, n_items), rnd.randint(0, n_items))
            if rnd.random() < 0.05 and r:
                r.append(r[0])
            rows.append(r)
        txt = "\n".join(" ".join(map(str, r)) for r in rows) + ("\n" if rnd.random() < 0.5 else "")
        if rnd.random() < 0.2:
            txt += "\n\n"
        db = tmp_path / "db.txt"
        db.write_text(txt)
        sup = rnd.choice([0.1, 0.2, 0.25, 0.3, 0.4, 0.5])
        ref = tmp_path / "o.res"
        oracle_lib.mine(db, sup, threads=2, out=ref)
        exp = tf.load_results(str(ref))
        for eng in ENGINES:
            out = tmp_path / "m.res"
            subprocess.run([mock_engine, str(db), repr(sup), "--out", str(out)] + eng, check=True,
                           capture_output=True)
            assert tf.load_results(str(out)) == exp, (it, eng, sup)


def test_parallel_parser_matches_one_pass(mock_engine, tmp_path):
    """The reader cut into pieces at '\\n' bytes gives the rows of one
    sequential pass: empty lines, '-' numbers, CR bytes, separators of any
    kind, a last line with or without '\\n', replicated or not."""
    rnd = random.Random(11)
    keys = ["transNum", "dimension", "itemIDmax", "dimensionReduced", "transNumReduced", "MFIsNumber",
            "transLengthMean", "transLengthReducedMean"]
    for it in range(40):
        lines = []
        for _ in range(rnd.randint(0, 60)):
            toks = [str(rnd.randint(0, 9)) for _ in range(rnd.randint(0, 8))]
            if rnd.random() < 0.2:
                toks.insert(rnd.randint(0, len(toks)), "-" + str(rnd.randint(1, 9)))
            sep = rnd.choice([" ", "  ", "\t", ",", " \r"])
            lines.append(sep.join(toks) + (" \r" if rnd.random() < 0.2 else ""))
        txt = "\n".join(lines) + rnd.choice(["", "\n", "\n\n", "  "])
        db = tmp_path / "db.txt"
        db.write_text(txt)
        rep = rnd.choice([[], ["--replicate", "3"]])
        outs = []
        for threads in (1, 2, 5, 17):
            out = tmp_path / ("p%d.res" % threads)
            r = subprocess.run([mock_engine, str(db), "0.3", "--out", str(out), "--engine", "wave",
                                "--parse-threads", str(threads)] + rep, capture_output=True, text=True)
            outs.append((r.returncode, tf.load_results(str(out)) if r.returncode == 0 else None,
                         {k: tf.load_header(str(out)).get(k) for k in keys} if r.returncode == 0 else None))
        assert all(o == outs[0] for o in outs), (it, txt[:200], rep, outs)


def test_device_preprocessing_flow_matches_host(mock_engine, tmp_path):
    """--prep gpu (counts, reduction and bitmaps through tfim_prep_begin /
    tfim_prep_finish, here the CPU stand-in) gives the host pass's header and
    output on databases with empty lines, repeated ids and item 0."""
    rnd = random.Random(5)
    keys = ["transNum", "dimension", "itemIDmin", "itemIDmax", "dimensionReduced", "transNumReduced", "MFIsNumber",
            "transLengthMin", "transLengthMax", "transLengthMean", "transLengthReducedMin", "transLengthReducedMax",
            "transLengthReducedMean"]
    for it in range(40):
        n_items = rnd.randint(2, 12)
        rows = []
        for _ in range(rnd.randint(0, 40)):
            r = rnd.sample(range(0, n_items), rnd.randint(0, n_items))
            if rnd.random() < 0.1 and r:
                r.append(r[-1])
            rows.append(r)
        txt = "\n".join(" ".join(map(str, r)) for r in rows) + rnd.choice(["", "\n", "\n\n"])
        db = tmp_path / "db.txt"
        db.write_text(txt)
        sup = rnd.choice([0.05, 0.2, 0.3, 0.5, 0.9])
        res = []
        for prep in ("host", "gpu"):
            out = tmp_path / ("%s.res" % prep)
            r = subprocess.run([mock_engine, str(db), repr(sup), "--out", str(out), "--prep", prep],
                               capture_output=True, text=True)
            res.append((r.returncode, tf.load_results(str(out)) if r.returncode == 0 else None,
                        {k: tf.load_header(str(out)).get(k) for k in keys} if r.returncode == 0 else None))
        assert res[0] == res[1], (it, txt[:200], sup, res)


def test_ids_past_the_limit_are_refused(mock_engine, tmp_path):
    """Ids above 2,147,483,645 (INT32_MAX - 2: per-id tables hold id_max + 2
    int32 entries) make the load fail with a message on both preprocessing
    paths instead of wrapping or looping."""
    for big in ("2147483646", "2147483647", "99999999999"):
        db = tmp_path / "big.txt"
        db.write_text("1 2 3\n2 %s\n" % big)
        for prep in ("host", "gpu"):
            r = subprocess.run([mock_engine, str(db), "0.5", "--out", str(tmp_path / "x.res"), "--prep", prep],
                               capture_output=True, text=True, timeout=60)
            assert r.returncode != 0, (big, prep, r.stdout, r.stderr)
            assert "exceeds 2147483645" in (r.stdout + r.stderr), (big, prep, r.stdout, r.stderr)
