"""The engines at the synthetic workloads' full sizes, against golden MFI
sets the oracle computed once (tests/golden/make_synthetic_fixtures.py):
clickstream (990,000 transactions, 1,758 survivors, support 0.1 %), dense
replicated (49,046-transaction base x256 = 12,555,776 transactions, support
0.85), gene (1,100 x 20,000, support 0.12) and dense small.  Each run's MFI
set and frequencies must match the fixture's digest exactly."""
import json
import os
import subprocess

import pytest

import tensorfim_amd as tf
from conftest import mfi_digest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "synthetic")

RUNS = {
    "clickstream": [("mafia", []), ("mafia", ["--learned"]), ("mafia", ["--compl", "off"]), ("wave", []),
                    ("plain", ["--workers", "8"])],
    "dense-replicated": [("mafia", []), ("mafia", ["--learned"]), ("mafia", ["--rows", "lists"]),
                         ("mafia", ["--graphs", "1"]), ("wave", []),
                         ("wave", ["--wave-kernel", "4"]), ("wave", ["--wave-kernel", "4", "--mfma-format", "int8"]),
                         ("plain", ["--workers", "8"])],
    "gene": [("mafia", []), ("mafia", ["--learned"]), ("mafia", ["--compl", "off"])],
    "dense-small": [("mafia", []), ("mafia", ["--learned"]), ("plain", [])],
}


@pytest.mark.parametrize("name", sorted(RUNS))
def test_full_size_matches_golden(tmp_path, name):
    with open(os.path.join(FIX, name + ".json")) as f:
        fx = json.load(f)
    data = tmp_path / (name + ".txt")
    subprocess.run([tf.binary("tfim_gen"), fx["workload"], str(data)], check=True)
    for engine, extra in RUNS[name]:
        if fx["replicate"] > 1:
            extra = extra + ["--replicate", str(fx["replicate"])]
        out = tmp_path / ("%s-%s.res" % (engine, "_".join(x.strip("-") for x in extra)))
        print("%s %s %s" % (name, engine, extra), flush=True)
        path, stats = tf.mine(str(data), fx["support"], engine=engine, out=str(out), extra=extra, timeout=300)
        got = tf.load_results(path)
        print("  mfis=%d rows=%s computation=%s" % (len(got), stats.get("rows"), stats.get("computation_time(s)")),
              flush=True)
        assert len(got) == fx["n_mfi"], (engine, extra)
        assert sum(got.values()) == fx["sum_freq"], (engine, extra)
        assert mfi_digest(got) == fx["digest"], (engine, extra)
