"""Mixed launches (hnswgpu_search_tasks_device): queries with their own k, ef
and filter share one launch through the device work queue, and each result is
the result of the uniform search_batch on that query, k, ef and filter -- ids,
distance bits and counts."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rows_equal(a, b):
    return np.array_equal(a["label"], b["label"]) and np.array_equal(a["distance"].view(np.uint32),
                                                                      b["distance"].view(np.uint32))


@pytest.fixture(scope="module")
def mixed_index(lib):
    import workloads

    X = workloads.mixture_128(20_000, seed=21)
    Q = workloads.mixture_128(500, seed=1021)
    ix = lib.Index(128, "l2")
    ix.add(X)
    return ix, X, Q


def check_against_uniform(dev, Q, ks, efs, fids, filters, storage=-1):
    got = dev.search_mixed(Q, ks, efs, fids, filters, storage=storage)
    groups = 0
    for k, ef, fid in sorted(set(zip(ks.tolist(), efs.tolist(), fids.tolist()))):
        idx = np.nonzero((ks == k) & (efs == ef) & (fids == fid))[0]
        want, wc = dev.search_rows(Q[idx], int(k), int(ef), None if fid < 0 else filters[fid])
        for j, i in enumerate(idx):
            rows, c = got[i]
            assert rows_equal(rows, want[j]), (k, ef, fid, i)
            assert c == wc[j]
        groups += 1
    return groups


def test_mixed_k_ef_filter_equal_uniform(lib, mixed_index):
    ix, X, Q = mixed_index
    dev = ix.device()
    rng = np.random.default_rng(0)
    n = Q.shape[0]
    ks = rng.choice([1, 10, 100], n)
    efs = rng.choice([0, 16, 64, 100, 200, 400, 700], n)  # 0 = ef_search; 700 > the register capacity
    fids = rng.choice([-1, 0, 1], n)
    filters = [rng.random(X.shape[0]) < 0.1, rng.random(X.shape[0]) < 0.5]
    assert check_against_uniform(dev, Q, ks, efs, fids, filters) > 30


def test_mixed_with_deletions_and_bf16(lib, mixed_index):
    import workloads

    X = workloads.mixture_128(8_000, seed=22)
    Q = workloads.mixture_128(300, seed=1022)
    ix = lib.Index(128, "l2")
    ix.add(X)
    for label in range(0, 8_000, 97):
        ix.mark_deleted(label)
    dev = ix.device()
    rng = np.random.default_rng(1)
    n = Q.shape[0]
    ks = rng.choice([5, 10], n)
    efs = rng.choice([10, 64, 128, 300], n)
    fids = rng.choice([-1, 0], n)
    filters = [rng.random(X.shape[0]) < 0.3]
    check_against_uniform(dev, Q, ks, efs, fids, filters)
    dev.set_option("storage", 1)
    check_against_uniform(dev, Q, ks, efs, fids, filters)
    dev.set_option("storage", 0)
    # An explicit storage argument traverses that copy without changing the option.
    got = dev.search_mixed(Q[:20], ks[:20], efs[:20], fids[:20], filters, storage=1)
    dev.set_option("storage", 1)
    for i in range(20):
        want, wc = dev.search_rows(Q[i:i + 1], int(ks[i]), int(efs[i]), None if fids[i] < 0 else filters[0])
        assert rows_equal(got[i][0], want[0]) and got[i][1] == wc[0]
    dev.set_option("storage", 0)


def test_mixed_deletions_without_filters_equal_oracle(lib, oracle, tmp_path):
    """Unfiltered mixed launches on an index with soft deletions take the
    per-query deletions kernel, whose ef classes nest (64 .. 512): with
    deleted_path 2 (side list of admitted deleted nodes at any count) and 0
    (the filtered kernel) every row equals the oracle's in the AVX-512 order
    (src/search_layer.cpp:58-134, deleted nodes are traversed but never
    admitted to the result)."""
    import workloads

    X = workloads.mixture_128(10_000, seed=24)
    Q = workloads.mixture_128(400, seed=1024)
    ix = lib.Index(128, "l2")
    ix.add(X)
    ix.save(tmp_path / "mixdel.idx")
    orc = oracle.load(tmp_path / "mixdel.idx")
    for label in range(5, 10_000, 89):
        ix.mark_deleted(label)
        orc.mark_deleted(label)
    dev = ix.device()
    rng = np.random.default_rng(4)
    n = Q.shape[0]
    ks = rng.choice([1, 10, 100], n)
    efs = rng.choice([10, 40, 100, 150, 250, 300, 500], n)
    fids = np.full(n, -1)
    want = {}
    for k, ef in sorted(set(zip(ks.tolist(), efs.tolist()))):
        idx = np.nonzero((ks == k) & (efs == ef))[0]
        want[(k, ef)] = (idx, orc.search(Q[idx], int(k), int(ef), order="avx512"))
    assert len(want) > 15
    for path in (2, 0):
        dev.set_option("deleted_path", path)
        got = dev.search_mixed(Q, ks, efs, fids, [])
        for (k, ef), (idx, w) in want.items():
            for j, i in enumerate(idx):
                rows, c = got[i]
                assert rows_equal(rows[None], w[j:j + 1]), (path, k, ef, i)
                assert c == int((w[j]["label"] != lib.INVALID_LABEL).sum()), (path, k, ef, i)
                assert not np.isin(rows["label"][:c].astype(np.int64), np.arange(5, 10_000, 89)).any()
    dev.set_option("deleted_path", 1)


def test_mixed_rejects_bad_tasks(lib, mixed_index):
    ix, X, Q = mixed_index
    dev = ix.device()
    with pytest.raises(ValueError):
        dev.search_mixed(Q[:2], [10, 0], [10, 10])
    with pytest.raises(ValueError):
        dev.search_mixed(Q[:2], [10, 10], [10, 10], [0, 3], [np.ones(X.shape[0], bool)])
    with pytest.raises(ValueError):
        dev.search_mixed(Q[:2], [10, 10], [10, 10], storage=2)  # fp8 copy never built on this index
