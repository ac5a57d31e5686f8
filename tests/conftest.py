"""Shared fixtures. GPU tests are marked ``gpu``; everything else runs on CPU."""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
sys.path.insert(0, str(ROOT / "hnsw-engine_amd"))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def golden():
    return np.load(GOLDEN / "golden.npz")


@pytest.fixture(scope="session")
def oracle():
    from oracle import Oracle

    return Oracle()


@pytest.fixture(scope="session")
def reference():
    from oracle import REF_SO, Reference

    if not REF_SO.exists():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    return Reference()


@pytest.fixture(scope="session")
def lib():
    import hnswgpu

    return hnswgpu


def case_names(golden):
    return sorted({k.split("/")[0] for k in golden.files})
