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


# ---------------------------------------------------------------------------
# MFMA prefilter (k <= 32, padded dim <= 256): certified queries come from the
# re-ranked shortlist, the others from the scan; rows must equal the scan's.
# ---------------------------------------------------------------------------

def both_paths(bf, Q, k, filter=None):
    bf.set_option("force_scan", 1)
    scan = bf.search_rows(Q, k, filter=filter)
    assert bf.get_option("last_mfma") == 0
    bf.set_option("force_scan", 0)
    fast = bf.search_rows(Q, k, filter=filter)
    return scan, fast


@pytest.mark.parametrize("metric,dim", [("l2", 128), ("ip", 100), ("cosine", 100), ("l2", 200), ("l2", 3)])
def test_mfma_prefilter_equals_scan(lib, metric, dim):
    rng = np.random.default_rng(dim)
    n, nq = 40_003, 333  # neither a multiple of the tile nor of the query block
    X = rng.standard_normal((n, dim)).astype(np.float32)
    Q = rng.standard_normal((nq, dim)).astype(np.float32)
    bf = lib.BruteForceIndex(dim, metric)
    bf.add(X, rng.permutation(n).astype(np.uint64) * 3 + 7)
    for k in (1, 10, 32):
        scan, fast = both_paths(bf, Q, k)
        assert bf.get_option("last_mfma") == 1
        assert rows_equal(fast, scan), (metric, dim, k, bf.get_option("last_redo"))
        # The shortlist certifies nearly every query on continuous data.
        assert bf.get_option("last_redo") <= nq // 20, bf.get_option("last_redo")


def test_mfma_prefilter_integer_mixture_and_chunks(lib, oracle):
    import workloads

    X = workloads.mixture_128(60_000, seed=4)
    Q = workloads.mixture_128(500, seed=1004)
    bf = lib.BruteForceIndex(128, "l2")
    bf.add(X)
    scan, _ = both_paths(bf, Q, 10)
    for variant in (0, 1):  # register-staged and global_load_lds tiles
        bf.set_option("mfma_variant", variant)
        for rows in (0, 2048, 64, 60_032):
            bf.set_option("chunk_rows", rows)
            fast = bf.search_rows(Q, 10)
            assert bf.get_option("last_mfma") == 1
            assert rows_equal(fast, scan), (variant, rows, bf.get_option("last_chunks"), bf.get_option("last_redo"))
    bf.set_option("chunk_rows", 0)
    bf.set_option("mfma_variant", 1)
    # Against the oracle's scan (AVX-512 order) on a few queries.
    want = oracle.brute_force(X, Q[:16], 10, 0, order="avx512")
    assert np.array_equal(scan["label"][:16].astype(np.int64), want["label"].astype(np.int64))
    assert np.array_equal(scan["distance"][:16].view(np.uint32), want["distance"].view(np.uint32))


def test_mfma_prefilter_with_filter(lib):
    rng = np.random.default_rng(11)
    n = 30_000
    X = rng.uniform(0, 1, (n, 64)).astype(np.float32)
    Q = rng.uniform(0, 1, (200, 64)).astype(np.float32)
    labels = np.arange(n, dtype=np.uint64) * 2
    bf = lib.BruteForceIndex(64, "l2")
    bf.add(X, labels)
    size = int(labels.max()) + 1
    allowed = rng.random(size) < 0.3
    words = np.zeros((size + 63) // 64, np.uint64)
    for i in np.nonzero(allowed)[0]:
        words[i >> 6] |= np.uint64(1) << np.uint64(i & 63)
    scan, fast = both_paths(bf, Q, 16, filter=(words, size))
    assert bf.get_option("last_mfma") == 1
    assert rows_equal(fast, scan)
    assert allowed[fast["label"].astype(np.int64)].all()
    # Filters that leave fewer than k rows: the lists hold every allowed row.
    few = np.zeros_like(words)
    few[0] = np.uint64(0b1010100)
    scan, fast = both_paths(bf, Q[:5], 10, filter=(few, size))
    assert rows_equal(fast, scan)
    assert (fast["label"][:, 3:] == lib.INVALID_LABEL).all()


def test_mfma_prefilter_falls_back_on_ties(lib):
    # 400 copies of one vector: every list fills with the same value, no
    # certificate is possible, and the scan orders the copies by label.
    rng = np.random.default_rng(5)
    X = rng.uniform(0, 1, (20_000, 32)).astype(np.float32)
    X[rng.choice(20_000, 400, replace=False)] = X[0]
    bf = lib.BruteForceIndex(32, "l2")
    bf.add(X, rng.permutation(20_000).astype(np.uint64))
    Q = np.vstack([X[:1], rng.uniform(0, 1, (63, 32)).astype(np.float32)])
    scan, fast = both_paths(bf, Q, 32)
    assert bf.get_option("last_redo") >= 1
    assert rows_equal(fast, scan)
    assert (fast["distance"][0] == 0).all()
    assert np.all(np.diff(fast["label"][0].astype(np.int64)) > 0)


def test_brute_force_options_are_checked(lib):
    bf = lib.BruteForceIndex(8, "l2")
    with pytest.raises(ValueError):
        bf.set_option("no_such_option", 1)
    with pytest.raises(ValueError):
        bf.set_option("last_redo", 1)
    bf.set_option("chunk_rows", 4096)
    assert bf.get_option("chunk_rows") == 4096
