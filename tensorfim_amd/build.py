"""Build script for the native parts of tensorfim_amd.

    python -m tensorfim_amd.build        (or __graft_entry__.build())

Compiles, in-tree:
  tensorfim_amd/lib/libtfim.so   HIP kernels + C ABI (include/tfim.h), gfx950
  tensorfim_amd/bin/tfim_mine    C++ engine (reference CLI), links libtfim.so
  tensorfim_amd/bin/tfim_gen     deterministic synthetic-dataset generator
  tensorfim_amd/bin/tfim_bench   isolated support-count microbench
"""
import os
import shlex
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "tensorfim_amd")
CSRC = os.path.join(PKG, "csrc")
LIB_DIR = os.path.join(PKG, "lib")
BIN_DIR = os.path.join(PKG, "bin")
OBJ_DIR = os.path.join(PKG, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17",
    # packed FP32 ops are disabled for every device compile on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
INCLUDES = ["-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
LIB_SOURCES = ["tfim_runtime.hip", "tfim_mlp.hip", "kernels_packed.hip", "kernels_mfma.hip", "kernels_mlp.hip",
               "kernels_prep.hip"]
CXX = os.environ.get("CXX", "g++")
CXX_FLAGS = ["-O3", "-std=c++17", "-march=x86-64-v2", "-Wall", "-Wno-unused-result"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n%s\n%s%s" % (" ".join(shlex.quote(c) for c in cmd), r.stdout, r.stderr))
    return r


def build_lib():
    os.makedirs(LIB_DIR, exist_ok=True)
    os.makedirs(OBJ_DIR, exist_ok=True)
    objs = []
    cmds = []
    for src in LIB_SOURCES:
        obj = os.path.join(OBJ_DIR, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC] + HIP_FLAGS + INCLUDES + ["-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(8, len(cmds))) as ex:
        list(ex.map(_run, cmds))
    out = os.path.join(LIB_DIR, "libtfim.so")
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC"] + objs + ["-o", tmp])
    os.replace(tmp, out)
    return out


def build_tools():
    os.makedirs(BIN_DIR, exist_ok=True)
    built = []
    eng = os.path.join(CSRC, "engine")
    rpath = "-Wl,-rpath,$ORIGIN/../lib"
    libs = ["-L" + LIB_DIR, "-ltfim", rpath, "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-lpthread"]
    targets = []
    if os.path.isdir(eng):
        srcs = sorted(os.path.join(eng, f) for f in os.listdir(eng) if f.endswith(".cpp") and not f.startswith("main_"))
        for main in sorted(f for f in os.listdir(eng) if f.startswith("main_") and f.endswith(".cpp")):
            name = "tfim_" + main[len("main_"):-len(".cpp")]
            targets.append((name, srcs + [os.path.join(eng, main)]))
    cmds = []
    for name, srcs in targets:
        out = os.path.join(BIN_DIR, name)
        cmds.append([CXX] + CXX_FLAGS + INCLUDES + ["-I" + eng] + srcs + ["-o", out + ".tmp"] + libs)
        built.append(out)
    with ThreadPoolExecutor(max_workers=max(1, min(8, len(cmds)))) as ex:
        list(ex.map(_run, cmds))
    for out in built:
        os.replace(out + ".tmp", out)
    return built


def build():
    lib = build_lib()
    tools = build_tools()
    return [lib] + tools


if __name__ == "__main__":
    for p in build():
        print(p)
    sys.exit(0)
