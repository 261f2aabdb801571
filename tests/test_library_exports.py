"""CPU: libtfim.so loads and exports every entry point include/tfim.h
declares (no compute calls: there is no GPU in this container)."""
import ctypes
import os
import re

import tensorfim_amd as tf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "tfim.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char|int64_t)\s*\**\s*(tfim_\w+)\s*\(", text, re.M)))


def test_header_declares_what_the_binding_uses():
    decl = declared_symbols()
    assert len(decl) >= 20
    assert set(tf.EXPORTED_SYMBOLS) <= set(decl)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(tf.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name
    tf.load_library()     # signatures bind
