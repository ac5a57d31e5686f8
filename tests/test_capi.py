"""The C ABI on CPU: the library loads and exports every function that
include/hnswgpu.h declares (no compute calls here)."""

from __future__ import annotations

import ctypes
import re

from conftest import ROOT


def declared_functions():
    text = (ROOT / "include" / "hnswgpu.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:hnswgpu|hnsw)_[a-z_0-9]+)\s*\(", text)))


def test_every_declared_symbol_is_exported(lib):
    names = declared_functions()
    assert len(names) >= 20
    so = ctypes.CDLL(str(lib.LIB_PATH))
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_neighbor_layout_is_16_bytes(lib):
    assert lib.NEIGHBOR.itemsize == 16


def test_last_error_is_thread_local_text(lib):
    import pytest

    with pytest.raises(ValueError, match="unknown metric"):
        lib.Index(4, "hamming")
