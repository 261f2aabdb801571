"""Copies the reference's own test data into tests/golden/ as fixtures.

The reference checks its engines by set-comparing each
`<dataset>-<thr>=Results.txt` with data/gold_results/ (scripts/verify_all.py,
scripts/run_all_configs.py).  Of its 15 gold files, 6 have their dataset in
the checkout (data/small/); those datasets (gzip -9) and golds are the
fixtures.  Run once, in a container that has the reference checkout:

    python tests/golden/make_fixtures.py /path/to/reference
"""
import gzip
import os
import shutil
import sys

PAIRS = [("chess.txt", "0.800000"), ("mushrooms.txt", "0.300000"), ("retail.txt", "0.010000"),
         ("T10I4D100K.txt", "0.005000"), ("connect.txt", "0.900000"), ("connect.txt", "0.850000")]


def main(ref):
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fimi")
    os.makedirs(here, exist_ok=True)
    for ds, thr in PAIRS:
        src = os.path.join(ref, "data", "small", ds)
        dst = os.path.join(here, ds + ".gz")
        if not os.path.exists(dst):
            with open(src, "rb") as f, gzip.GzipFile(dst, "wb", compresslevel=9, mtime=0) as g:
                shutil.copyfileobj(f, g)
        gold = "%s-%s=Results.txt" % (ds, thr)
        shutil.copyfile(os.path.join(ref, "data", "gold_results", gold), os.path.join(here, gold))
    print("fixtures in", here)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
