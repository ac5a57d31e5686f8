"""GPU parity of BruteForceIndex: equal to the reference's golden rows
(AVX-512 order, best k by (distance, label), filter applied first) and to an
FP64 scan within fp32 rounding."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import case_names

pytestmark = pytest.mark.gpu


def rows_equal(got, want):
    return np.array_equal(got["label"], want["label"]) and np.array_equal(
        got["distance"].view(np.uint32), want["distance"].view(np.uint32))


def test_brute_force_equals_reference_golden(golden, lib):
    for name in case_names(golden):
        X, Q, lab = golden[f"{name}/data"], golden[f"{name}/queries"], golden[f"{name}/labels"]
        metric = int(golden[f"{name}/meta"][0])
        bf = lib.BruteForceIndex(X.shape[1], lib.METRIC_NAMES[metric])
        bf.add(X, lab)
        k = min(10, X.shape[0])
        assert rows_equal(bf.search_rows(Q, k), golden[f"{name}/avx512/brute/k10"]), name
        filt = (golden[f"{name}/allow_words"], int(golden[f"{name}/allow_size"][0]))
        assert rows_equal(bf.search_rows(Q, k, filter=filt), golden[f"{name}/avx512/brute_filter/k10"]), name


def test_brute_force_fp64_and_padding(lib):
    rng = np.random.default_rng(3)
    X = rng.uniform(-1, 1, (500, 19)).astype(np.float32)
    q = rng.uniform(-1, 1, (3, 19)).astype(np.float32)
    bf = lib.BruteForceIndex(19, "l2")
    bf.add(X)
    got = bf.search_rows(q, 10)
    d = ((q[:, None, :].astype(np.float64) - X[None]) ** 2).sum(-1)
    assert np.array_equal(got["label"], np.argsort(d, axis=1, kind="stable")[:, :10])
    np.testing.assert_allclose(got["distance"], np.sort(d, axis=1)[:, :10], atol=1e-4)
    small = lib.BruteForceIndex(4, "ip")
    small.add(X[:3, :4])
    got = small.search_rows(X[:1, :4], 10)
    assert (got["label"][0, 3:] == lib.INVALID_LABEL).all() and np.isinf(got["distance"][0, 3:]).all()
