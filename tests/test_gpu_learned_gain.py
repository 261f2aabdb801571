"""What the trained network gains over the search without it.

The MAFIA-style engine without a model runs the classical head-union-tail
lookahead at every node with a tail of two or more (``--gate always``).  With
``--learned`` the MLP is trained online on the GPU from the engine's own
lookahead outcomes and decides which lookaheads to run (an analytic
independence estimate proposes, the model adds, a small exploration share
keeps labels coming; the gate threshold and the exploration share calibrate
themselves from the outcomes of the lookaheads they add).  Every decision is confirmed by exact counts, so the MFI set
must not change; what changes is the number of counted rows, mostly the
lookaheads that fail.

The dense replicated workload (49,046 base transactions x 16, support 0.85)
is the measured case.  The analytic rule alone (``--gate indep``) is run
beside it and reported, so the model's share of the gain is visible.
"""
import subprocess

import pytest

import tensorfim_amd as tf

pytestmark = pytest.mark.gpu


def _run(data, tmp_path, name, extra):
    out = str(tmp_path / (name + ".res"))
    path, st = tf.mine(str(data), 0.85, engine="mafia", out=out, extra=["--replicate", "16"] + extra, timeout=600)
    keys = ("MFIsNumber", "rows", "hut_rows", "hut_success", "model_lookaheads", "mlp_train_steps",
            "computation_time(s)", "learn_time(s)")
    print(name, {k: st.get(k) for k in keys}, flush=True)
    return tf.load_results(path), st


def test_learned_gate_gain_dense_replicated(tmp_path):
    data = tmp_path / "dense.txt"
    subprocess.run([tf.binary("tfim_gen"), "dense-replicated", str(data)], check=True)
    classic, sc = _run(data, tmp_path, "always", ["--gate", "always"])
    learned, sl = _run(data, tmp_path, "learned", ["--learned"])
    rule, sr = _run(data, tmp_path, "indep", ["--gate", "indep"])
    assert learned == classic == rule
    assert int(sl["mlp_train_steps"]) > 0, "the network was never trained"
    failed = lambda st: int(st["hut_rows"]) - int(st["hut_success"])
    # the trained gate runs far fewer lookaheads that fail ...
    assert failed(sl) * 4 <= failed(sc), (failed(sl), failed(sc))
    # ... keeps nearly every lookahead that succeeds ...
    assert int(sl["hut_success"]) >= 0.9 * int(sc["hut_success"])
    # ... and counts fewer rows in total
    assert int(sl["rows"]) < int(sc["rows"])


def test_learned_policy_trains_on_gene(tmp_path):
    """The wide gene-expression-like workload (1,100 x 20,000, support 0.12)
    has few lookahead decisions and none of them succeeds: the network must
    still be trained online (gate labels from the warm-up sample, ranker
    labels from sampled nodes' children), and the output must equal the
    search without it."""
    data = tmp_path / "gene.txt"
    subprocess.run([tf.binary("tfim_gen"), "gene", str(data)], check=True)
    res = {}
    for name, extra in (("learned", ["--learned"]), ("indep", []), ("always", ["--gate", "always"])):
        out = str(tmp_path / (name + ".res"))
        path, st = tf.mine(str(data), 0.12, engine="mafia", out=out, extra=extra, timeout=600)
        print(name, {k: st.get(k) for k in ("MFIsNumber", "rows", "hut_rows", "hut_success", "mlp_train_steps",
                                            "model_lookaheads", "computation_time(s)")}, flush=True)
        res[name] = (tf.load_results(path), st)
    assert res["learned"][0] == res["indep"][0] == res["always"][0]
    assert int(res["learned"][1]["mlp_train_steps"]) > 0, "the network was never trained"
    # the trained gate looks ahead far less than the classical rule, whose
    # lookaheads all fail here
    assert int(res["learned"][1]["hut_rows"]) * 2 <= int(res["always"][1]["hut_rows"])
