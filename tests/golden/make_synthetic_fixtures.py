"""Golden MFI sets of the synthetic workloads at their full sizes.

No reference gold exists for the synthetic workloads, so the oracle's
reference DFS (oracle/, pinned by the reference's own gold files) is the
golden source.  For each workload the generator (tfim_gen) writes the
transaction file, the oracle mines it and the fixture keeps the MFI count,
the sum of frequencies and a SHA-256 digest of the canonical listing (one
line "sorted ids:frequency" per MFI, lines sorted).  The dense replicated
workload is mined on its 49,046-transaction base and every frequency scaled
by 256 (data/make_replica.py concatenates copies; both thresholds are
non-integers, so the sets coincide).  Takes about ten minutes on 8 cores
(the clickstream workload dominates).  Run after building:

    python tests/golden/make_synthetic_fixtures.py
"""
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle_lib  # noqa: E402
import tensorfim_amd as tf  # noqa: E402
from conftest import mfi_digest  # noqa: E402

# (name, generator workload, support, replication applied to the base)
WORKLOADS = [("dense-small", "dense-small", 0.6, 1), ("clickstream", "clickstream", 0.001, 1),
             ("dense-replicated", "dense-replicated", 0.85, 256), ("gene", "gene", 0.12, 1)]


def main():
    out_dir = os.path.join(HERE, "synthetic")
    os.makedirs(out_dir, exist_ok=True)
    threads = os.cpu_count() or 8
    for name, gen, support, rep in WORKLOADS:
        with tempfile.TemporaryDirectory() as d:
            data = os.path.join(d, name + ".txt")
            subprocess.run([tf.binary("tfim_gen"), gen, data], check=True)
            res_path = os.path.join(d, "oracle.res")
            stats = oracle_lib.mine(data, support, threads=threads, out=res_path, timeout=3600)
            res = {s: f * rep for s, f in tf.load_results(res_path).items()}
            hdr = tf.load_header(res_path)
        fx = {"workload": gen, "support": support, "replicate": rep, "n_mfi": len(res),
              "sum_freq": sum(res.values()), "digest": mfi_digest(res), "oracle_nodes": int(stats["nodes"]),
              "base_header": {k: hdr[k] for k in ("transNum", "dimensionReduced", "transNumReduced", "MFIsNumber")}}
        with open(os.path.join(out_dir, name + ".json"), "w") as f:
            json.dump(fx, f, indent=1, sort_keys=True)
        print(name, fx["n_mfi"], fx["digest"][:16], flush=True)


if __name__ == "__main__":
    main()
