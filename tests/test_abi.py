"""The C-ABI library loads and exports every entry point the header declares
(no device call: runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols(header="pnamp_gpu.h"):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(pnamp_[a-z_0-9]+)\s*\(", src)))


def test_header_symbols_exported():
    from sccm_amd import pnamp
    lib = ctypes.CDLL(pnamp.LIB_PATH)
    syms = declared_symbols() + declared_symbols("pnamp_mlp.h")
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_wrapper_binds_all_symbols():
    from sccm_amd import pnamp
    L = pnamp.load_library()
    for s in declared_symbols():
        assert getattr(L, s).restype is not None or s == "pnamp_destroy"


def test_surrogate_wrapper_binds_all_symbols():
    from sccm_amd import surrogate
    L = surrogate.load_mlp()
    for s in declared_symbols("pnamp_mlp.h"):
        assert getattr(L, s).restype is not None or s == "pnamp_mlp_destroy"
