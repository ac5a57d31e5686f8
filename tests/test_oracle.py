"""Pins the oracle (oracle/oracle.cpp) before anything is checked against it.

* distances: bit-exact against the reference's own kernels (compiled into
  oracle/_ref) in all three summation orders, exact on small integers (closed
  form), and different between orders on random data (so an order mix-up is
  caught);
* search: equal to the reference's golden vectors (tests/golden) in the order
  each index was built with, for plain / filtered / deleted / both;
* known answers: the 1-D line (tests/cpp/test_index.cpp:127-139 of the
  reference), ef >= n equals exact search, cosine closed form 1 - sqrt(0.5).
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import case_names

ISAS = ("scalar", "avx512")


def test_distance_orders_bit_exact_vs_reference_kernels(oracle, reference):
    rng = np.random.default_rng(0)
    a = rng.uniform(-1, 1, 1025).astype(np.float32)
    b = rng.uniform(-1, 1, 1025).astype(np.float32)
    import ctypes as C

    for isa in ("scalar", "avx2", "avx512"):
        reference.set_isa(isa)
        for dim in range(1, 1025):
            if isa == "avx2" and dim % 8:
                continue
            for off in (0, 1):
                pa, pb = a.ctypes.data + 4 * off, b.ctypes.data + 4 * off
                want_l2 = np.float32(reference.lib.ref_l2_sq(C.c_void_p(pa), C.c_void_p(pb), dim))
                want_dot = np.float32(reference.lib.ref_dot(C.c_void_p(pa), C.c_void_p(pb), dim))
                got_l2 = np.float32(oracle.l2(a[off:off + dim], b[off:off + dim], isa))
                got_dot = np.float32(oracle.dot(a[off:off + dim], b[off:off + dim], isa))
                assert got_l2.tobytes() == want_l2.tobytes(), (isa, dim, off)
                assert got_dot.tobytes() == want_dot.tobytes(), (isa, dim, off)
    reference.set_isa("avx512")


@pytest.mark.parametrize("order", ["scalar", "avx2", "avx512"])
def test_distance_exact_on_small_integers(oracle, order):
    # Closed form (tests/cpp/test_kernels.cpp:44-58 of the reference).
    for dim in (8, 16, 64, 96, 128, 960) if order == "avx2" else (1, 7, 8, 9, 15, 16, 17, 31, 33, 63, 64, 65, 100,
                                                                    128, 960):
        i = np.arange(dim)
        a = (i % 7 - 3).astype(np.float32)
        b = (i % 5 - 2).astype(np.float32)
        assert oracle.l2(a, b, order) == float(((a.astype(np.float64) - b) ** 2).sum())
        assert oracle.dot(a, b, order) == float((a.astype(np.float64) * b).sum())


def test_summation_orders_really_differ(oracle):
    rng = np.random.default_rng(1)
    differ = 0
    for _ in range(200):
        a = rng.standard_normal(128).astype(np.float32)
        b = rng.standard_normal(128).astype(np.float32)
        differ += oracle.l2(a, b, "scalar") != oracle.l2(a, b, "avx512")
    assert differ > 20


def _rows_equal(got, want):
    return np.array_equal(got["label"], want["label"]) and np.array_equal(
        got["distance"].view(np.uint32), want["distance"].view(np.uint32))


@pytest.mark.parametrize("isa", ISAS)
def test_oracle_search_equals_reference_golden(golden, oracle, isa):
    from conftest import GOLDEN

    for name in case_names(golden):
        idx = oracle.load(GOLDEN / f"{name}_{isa}.idx")
        Q = golden[f"{name}/queries"]
        filt = (golden[f"{name}/allow_words"], int(golden[f"{name}/allow_size"][0]))
        for key in golden.files:
            parts = key.split("/")
            if len(parts) != 4 or parts[0] != name or parts[1] != isa or parts[2] not in ("plain", "filter"):
                continue
            k, ef = (int(x[1:]) if x[0] == "k" else int(x[2:]) for x in parts[3].split("_"))
            got = idx.search(Q, k, ef, order=isa, filter=filt if parts[2] == "filter" else None)
            assert _rows_equal(got, golden[key]), key
        for lab in golden[f"{name}/deleted"].tolist():
            idx.mark_deleted(lab)
        assert _rows_equal(idx.search(Q, 10, 64, order=isa), golden[f"{name}/{isa}/deleted/k10_ef64"]), name
        assert _rows_equal(idx.search(Q, 10, 64, order=isa, filter=filt),
                           golden[f"{name}/{isa}/deleted_filter/k10_ef64"]), name


def test_oracle_line_known_answer(oracle):
    from conftest import GOLDEN

    for isa in ISAS:
        idx = oracle.load(GOLDEN / f"line_d1_{isa}.idx")
        got = idx.search(np.array([[42.2]], np.float32), 3, 50, order=isa)
        assert got["label"][0].tolist() == [42, 43, 41]


def test_oracle_large_ef_is_exact(golden, oracle):
    from conftest import GOLDEN

    X = golden["l2_d32/data"]
    Q = golden["l2_d32/queries"]
    idx = oracle.load(GOLDEN / "l2_d32_avx512.idx")
    got = idx.search(Q, 10, X.shape[0], order="avx512")
    d64 = ((Q[:, None, :].astype(np.float64) - X[None, :, :]) ** 2).sum(-1)
    want = np.argsort(d64, axis=1, kind="stable")[:, :10]
    assert np.array_equal(got["label"], want)
    np.testing.assert_allclose(got["distance"], np.take_along_axis(d64, want, 1), rtol=1e-5)


def test_oracle_brute_force_vs_golden_and_fp64(golden, oracle):
    for name in case_names(golden):
        X, Q, lab = golden[f"{name}/data"], golden[f"{name}/queries"], golden[f"{name}/labels"]
        metric = int(golden[f"{name}/meta"][0])
        k = min(10, X.shape[0])
        filt = (golden[f"{name}/allow_words"], int(golden[f"{name}/allow_size"][0]))
        for isa in ISAS:
            got = oracle.brute_force(X, Q, k, metric=metric, labels=lab, order=isa)
            assert _rows_equal(got, golden[f"{name}/{isa}/brute/k10"]), (name, isa)
            got = oracle.brute_force(X, Q, k, metric=metric, labels=lab, order=isa, filter=filt)
            assert _rows_equal(got, golden[f"{name}/{isa}/brute_filter/k10"]), (name, isa)
        if metric == 0:
            d64 = ((Q[:, None, :].astype(np.float64) - X[None, :, :]) ** 2).sum(-1)
            got = oracle.brute_force(X, Q, k, metric=0, labels=lab)
            np.testing.assert_allclose(got["distance"], np.sort(d64, axis=1)[:, :k], rtol=1e-5, atol=1e-5)


def test_oracle_cosine_closed_form(oracle):
    # tests/cpp/test_brute_force.cpp:34-47 of the reference.
    data = np.zeros((3, 8), np.float32)
    data[0, 0] = 1
    data[1, 1] = 5
    data[2, :2] = 3
    q = np.zeros((1, 8), np.float32)
    q[0, 1] = 100
    got = oracle.brute_force(data, q, 3, metric=2)
    assert got["label"][0].tolist()[:2] == [1, 2]
    assert abs(got["distance"][0, 0]) <= 1e-6
    assert abs(got["distance"][0, 1] - (1 - np.sqrt(0.5))) <= 1e-6
