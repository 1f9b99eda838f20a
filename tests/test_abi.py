"""The C-ABI library loads and exports every entry point include/*.h declares."""
import ctypes
import glob
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "origami_amd", "_lib", "liborigami_gpu.so")


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(og_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


@pytest.fixture(scope="module")
def lib_path():
    if not os.path.exists(LIB):
        from origami_amd import build
        build.build()
    return LIB


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("og_engine_create", "og_forces", "og_energies", "og_run", "og_swap_systems",
              "og_dynamics_setup", "og_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol(lib_path):
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], check=True,
                         capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = sorted(declared_functions() - exported)
    assert not missing, "declared but not exported: %s" % missing


def test_library_loads(lib_path):
    lib = ctypes.CDLL(lib_path)
    for n in declared_functions():
        assert hasattr(lib, n), n
    lib.og_last_error.restype = ctypes.c_char_p
    assert lib.og_last_error() is not None


def test_create_rejects_bad_input_without_gpu(lib_path):
    # argument validation happens before any device call
    lib = ctypes.CDLL(lib_path)
    lib.og_last_error.restype = ctypes.c_char_p
    out = ctypes.c_void_p()
    assert lib.og_engine_create(None, 0, ctypes.byref(out)) == -1
    assert b"at least one system" in lib.og_last_error()
