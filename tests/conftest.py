import gzip
import os
import shutil
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

FIMI = os.path.join(ROOT, "tests", "golden", "fimi")
GOLD_PAIRS = [("chess.txt", 0.8), ("mushrooms.txt", 0.3), ("retail.txt", 0.01),
              ("T10I4D100K.txt", 0.005), ("connect.txt", 0.9), ("connect.txt", 0.85)]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


@pytest.fixture(scope="session")
def fimi_dir(tmp_path_factory):
    """The gold datasets, decompressed into a temp dir (Results files land there)."""
    d = tmp_path_factory.mktemp("fimi")
    for ds in sorted({p[0] for p in GOLD_PAIRS}):
        with gzip.open(os.path.join(FIMI, ds + ".gz"), "rb") as f, open(d / ds, "wb") as g:
            shutil.copyfileobj(f, g)
    return d


def mfi_digest(results):
    """SHA-256 of a {frozenset(items): frequency} MFI set: one line
    "sorted ids:frequency" per MFI, lines sorted."""
    import hashlib
    lines = sorted(" ".join(map(str, sorted(s))) + ":" + str(f) for s, f in results.items())
    return hashlib.sha256("\n".join(lines).encode()).hexdigest()


def gold_path(ds, thr):
    return os.path.join(FIMI, "%s-%f=Results.txt" % (ds, thr))
