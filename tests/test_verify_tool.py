"""CPU: tfim_verify, the C++ restatement of the reference checker
(scripts/verify_all.py:13-23,40), against the Python comparison used by the
rest of the tests, on the gold Results files."""
import os
import shutil
import subprocess

import pytest

import tensorfim_amd as tf
from conftest import GOLD_PAIRS, gold_path


def _verify(*args):
    return subprocess.run([tf.binary("tfim_verify")] + list(args), capture_output=True, text=True)


@pytest.mark.parametrize("ds,thr", GOLD_PAIRS)
def test_gold_against_itself_and_counts(ds, thr):
    g = gold_path(ds, thr)
    r = _verify(g, g, "--header")
    assert r.returncode == 0, r.stdout
    assert "(%d MFIs)" % len(tf.load_results(g)) in r.stdout


def test_mismatches_are_found(tmp_path):
    ds, thr = GOLD_PAIRS[0]
    g = gold_path(ds, thr)
    lines = open(g, encoding="utf-8").read().split("\n")
    idx = next(i for i, ln in enumerate(lines) if "frequency:" in ln and "th:" in ln)
    ln = lines[idx]
    head, f = ln.rsplit("frequency:", 1)
    changed = lines[:idx] + [head + "frequency: %d" % (int(f) + 1)] + lines[idx + 1:]
    p = tmp_path / "changed.txt"
    p.write_text("\n".join(changed), encoding="utf-8")
    r = _verify(str(p), g)
    assert r.returncode == 1 and "[MISMATCH]" in r.stdout
    dropped = lines[:idx] + lines[idx + 1:]
    p.write_text("\n".join(dropped), encoding="utf-8")
    r = _verify(str(p), g)
    assert r.returncode == 1 and "(missing)" in r.stdout
    # reordered lines and extra spaces compare equal, as sets
    body = [x for x in lines if "frequency:" in x and "th:" in x]
    rest = [x for x in lines if x not in body]
    p.write_text("\n".join(rest + ["  " + x.replace(" ", "  ") for x in reversed(body)]), encoding="utf-8")
    assert _verify(str(p), g).returncode == 0


def test_directory_mode(tmp_path):
    res, gold = tmp_path / "res", tmp_path / "gold"
    res.mkdir()
    gold.mkdir()
    for ds, thr in GOLD_PAIRS[:3]:
        g = gold_path(ds, thr)
        shutil.copy(g, gold / os.path.basename(g))
        shutil.copy(g, res / os.path.basename(g))
    (res / "other.txt-0.500000=Results.txt").write_text("no gold for this one\n")
    r = _verify("--dir", str(res), str(gold), "--header")
    assert r.returncode == 0 and "3 identical, 0 mismatched" in r.stdout and "[no gold]" in r.stdout
