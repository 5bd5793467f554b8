"""The C-ABI library loads on a CPU-only host and exports every entry point
declared in include/ppl_amd.h (no compute call is made)."""
import ctypes
import os
import re

from conftest import ROOT, load_ppl


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "ppl_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ppl_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    ppl = load_ppl()
    L = ctypes.CDLL(ppl.LIB_PATH)
    names = declared_symbols()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_header():
    ppl = load_ppl()
    assert set(declared_symbols()) == set(ppl.SIGNATURES)


def test_last_error_without_gpu():
    ppl = load_ppl()
    L = ppl.lib()
    assert isinstance(L.ppl_last_error(), bytes)
    assert L.ppl_version() >= 1
