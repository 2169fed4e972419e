"""The C ABI library loads and exports every symbol include/wb_gpu.h declares
(no compute calls: this runs without a GPU)."""
import ctypes as C
import os
import re

import wolkenbase_amd as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "wb_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wb_[a-z_0-9]+)\s*\(", text)))


def test_header_symbols_exported():
    syms = declared_symbols()
    assert len(syms) >= 14
    L = C.CDLL(W.GPU_SO)
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_python_signatures_cover_header():
    assert set(declared_symbols()) <= set(W.SIGNATURES)


def test_null_arguments_rejected_without_gpu():
    L = W.lib()
    assert L.wb_size_fit(None, 0, None, None) == -1
    assert L.wb_tile_count(None, None, 1.0, None) == -1
