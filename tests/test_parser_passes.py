"""The parser's two passes agree (tensorfim_amd/csrc/engine/dataset.cpp).

The first pass only counts ids and rows per piece, branch-free; the second
parses.  The reader's rules they both follow are those of kernel_bmma.cu:
208-270 (a digit run is an id, '-' drops the digit run after it, any other
byte separates, every '\\n' ends a row).  Runs on the CPU.
"""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ENG = os.path.join(ROOT, "tensorfim_amd", "csrc", "engine")


def test_count_and_parse_passes_agree(tmp_path):
    exe = str(tmp_path / "parser_passes")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), "-I" + ENG,
                    os.path.join(ROOT, "tests", "native", "parser_passes.cpp"),
                    os.path.join(ROOT, "tests", "engine_mock", "mock_tfim.cpp"), "-o", exe, "-lpthread"], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "bad 0" in r.stdout
