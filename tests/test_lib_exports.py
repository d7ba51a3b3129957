"""The C-ABI library loads and exports every symbol the headers declare
(no compute calls: this runs without a GPU)."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "readdygpu_amd", "libCReaDDyGPU.so")


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in sorted(os.listdir(inc)):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b((?:rd|Java)_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        import __graft_entry__
        __graft_entry__.build()
    return C.CDLL(LIB)


def test_headers_declare_the_c_abi():
    names = declared_symbols()
    assert "rd_simulate" in names and "rd_create" in names
    assert len(names) >= 20


def test_every_declared_symbol_is_exported(lib):
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_bindings_cover_the_header(lib):
    from readdygpu_amd import _lib
    declared = {n for n in declared_symbols() if n.startswith("rd_")}
    assert declared <= set(_lib._SIGS), declared - set(_lib._SIGS)


def test_exported_symbols_are_extern_c(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    for n in declared_symbols():
        assert re.search(rf"\bT {n}$", out, flags=re.M), n
