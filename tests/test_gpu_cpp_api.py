"""The engine's C++ API (include/hnsw/*.hpp) from a C++ program: tests/cpp/
api_check, built by __graft_entry__.build() against libhnswgpu.so, checks the
behaviours of the reference's C++ tests (parameter errors, empty index, the
line known answer, k above the size, duplicate labels, identical vectors,
search == search_batch rows, BitsetFilter == FunctionFilter, concurrent
searches, soft deletes, an mmap-loaded copy, brute force == an FP64 scan). Its
search_batch rows on the index it wrote must equal the oracle's bit for bit."""

from __future__ import annotations

import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

BIN = ROOT / "tests" / "cpp" / "api_check"
ROW = np.dtype([("label", "<u8"), ("distance", "<f4"), ("pad", "<u4")])


def test_cpp_api_program_and_oracle_rows(oracle, tmp_path):
    assert BIN.exists(), "tests/cpp/api_check is not built (__graft_entry__.build() makes it)"
    print("running", BIN, flush=True)
    r = subprocess.run([str(BIN), str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    q = np.fromfile(tmp_path / "queries.f32", np.float32).reshape(200, 32)
    rows = np.fromfile(tmp_path / "rows.bin", ROW).reshape(200, 10)
    want = oracle.load(tmp_path / "api.idx").search(q, 10, 64, order="avx512")
    np.testing.assert_array_equal(rows["label"], want["label"])
    np.testing.assert_array_equal(rows["distance"].view(np.uint32), want["distance"].view(np.uint32))
