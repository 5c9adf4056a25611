"""The C-ABI library builds, loads and exports every symbol include/pdrchem.h
declares (no compute calls: this runs without a GPU)."""
import os
import re
import subprocess

import pytest

from pdrchem import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pdrchem.h")


def declared_symbols():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    funcs = set(re.findall(r"\b([a-z_][a-z0-9_]*)\s*\(", src)) - {"if", "sizeof"}
    funcs = {f for f in funcs if f.startswith("pdrchem_") or f == "calculate_abundances_"}
    globs = set(re.findall(r"extern\s+(?:double|int)\s+([a-z_0-9]+)\s*;", src))
    return funcs, globs


def test_library_exports_declared_symbols():
    assert os.path.exists(lib.LIB_PATH), "libpdrchem.so not built"
    out = subprocess.run(["nm", "-D", "--defined-only", lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    funcs, globs = declared_symbols()
    assert funcs == set(lib.EXPORTED_FUNCTIONS)
    assert globs == {g for g, _ in lib.EXPORTED_GLOBALS}
    missing = (funcs | globs) - exported
    assert not missing, missing


def test_library_loads_and_globals_default():
    L = lib.load()
    g = lib.get_globals()
    assert g["chemistry_module_mp_relative_abundance_tolerance_"] == 1e-8
    assert g["chemistry_module_mp_absolute_abundance_tolerance_"] == 1e-30
    assert g["maincode_module_mp_end_time_"] == 1e7
    assert lib.network_id("reduced") == 0
    assert L.pdrchem_network_id(12, 34) == -1


def test_globals_writable():
    lib.set_globals(end_time=5e6)
    assert lib.get_globals()["maincode_module_mp_end_time_"] == 5e6
    lib.set_globals(end_time=1e7)
