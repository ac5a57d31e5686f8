"""GPU parity of the graph search (through the C ABI).

The bar is exact: the GPU computes distances in the AVX-512 summation order,
so on an index the reference built under AVX-512 it must return the
reference's golden rows bit for bit, and on any index it must equal the
oracle run in the same order -- ids, order and fp32 distance bits.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import GOLDEN, case_names

pytestmark = pytest.mark.gpu


def rows_equal(got, want):
    return np.array_equal(got["label"], want["label"]) and np.array_equal(
        got["distance"].view(np.uint32), want["distance"].view(np.uint32))


def parse_k_ef(tag):
    k, ef = tag.split("_")
    return int(k[1:]), int(ef[2:])


def test_equals_reference_golden_rows(golden, lib):
    for name in case_names(golden):
        print("case", name, flush=True)
        ix = lib.Index.load(GOLDEN / f"{name}_avx512.idx")
        Q = golden[f"{name}/queries"]
        filt = (golden[f"{name}/allow_words"], int(golden[f"{name}/allow_size"][0]))
        for kind in ("plain", "filter"):
            for key in golden.files:
                if key.startswith(f"{name}/avx512/{kind}/"):
                    k, ef = parse_k_ef(key.split("/")[-1])
                    got = ix.search_rows(Q, k, ef, filter=filt if kind == "filter" else None)
                    assert rows_equal(got, golden[key]), key
        for lab in golden[f"{name}/deleted"].tolist():
            ix.mark_deleted(lab)
        assert rows_equal(ix.search_rows(Q, 10, 64), golden[f"{name}/avx512/deleted/k10_ef64"]), name
        assert rows_equal(ix.search_rows(Q, 10, 64, filter=filt), golden[f"{name}/avx512/deleted_filter/k10_ef64"])


def test_equals_oracle_on_scalar_built_indexes(golden, lib, oracle):
    for name in case_names(golden):
        print("case", name, flush=True)
        path = GOLDEN / f"{name}_scalar.idx"
        ix, orc = lib.Index.load(path), oracle.load(path)
        Q = golden[f"{name}/queries"]
        for k, ef in ((1, 1), (10, 10), (10, 64), (7, 300)):
            assert rows_equal(ix.search_rows(Q, k, ef), orc.search(Q, k, ef, order="avx512")), (name, k, ef)


def test_general_kernel_equals_fast_kernel(golden, lib):
    for name in ("l2_d32", "ties_d8", "cos_d100"):
        ix = lib.Index.load(GOLDEN / f"{name}_avx512.idx")
        Q = golden[f"{name}/queries"]
        filt = (golden[f"{name}/allow_words"], int(golden[f"{name}/allow_size"][0]))
        dev = ix.device()
        fast = ix.search_rows(Q, 10, 64, filter=filt)
        dev.set_option("force_spill", 1)
        slow = ix.search_rows(Q, 10, 64, filter=filt)
        dev.set_option("force_spill", 0)
        assert rows_equal(fast, slow), name


def test_visited_set_spills_to_hbm_and_stays_exact(golden, lib, oracle):
    path = GOLDEN / "l2_d32_scalar.idx"
    ix, orc = lib.Index.load(path), oracle.load(path)
    Q = golden["l2_d32/queries"]
    want = orc.search(Q, 10, 64, order="avx512")
    dev = ix.device()
    # A 256-word LDS hash (256 32-bit slots, or 512 16-bit ones): queries
    # continue in their HBM tables.
    for wide, ef in ((1, 64), (0, 200)):
        dev.set_option("wide_hash", wide)
        dev.set_option("hash_log2", 8)
        rows, _ = dev.search_rows(Q, 10, ef)
        assert dev.get_option("last_hbm_visited") > 0, wide
        assert dev.get_option("last_overflow") == 0
        assert rows_equal(rows, orc.search(Q, 10, ef, order="avx512")), wide
    dev.set_option("wide_hash", 1)
    # A 256-slot HBM table as well: the rest go to the general kernel.
    dev.set_option("visit_log2", 8)
    rows, _ = dev.search_rows(Q, 10, 64)
    assert dev.get_option("last_handoff_visited") > 0
    assert rows_equal(rows, want)
    dev.set_option("visit_log2", 15)
    dev.set_option("hash_log2", 0)
    for wide in (1, 0):
        dev.set_option("wide_hash", wide)
        rows, _ = dev.search_rows(Q, 10, 64)
        assert rows_equal(rows, want)


def test_uniform_20k_x_128_parity(lib, oracle, tmp_path):
    import workloads

    X = workloads.uniform(20000, 128, 0)
    Q = workloads.uniform(1000, 128, 1)
    ix = lib.Index(128, "l2", M=16, ef_construction=200)
    ix.add(X, num_threads=0)
    ix.save(tmp_path / "u.idx")
    orc = oracle.load(tmp_path / "u.idx")
    for ef in (10, 64, 128, 256, 512):
        print("ef", ef, flush=True)
        got = ix.search_rows(Q, 10, ef)
        assert rows_equal(got, orc.search(Q, 10, ef, order="avx512", threads=8)), ef
    # Uniform 128-d data is hard for graph search: recall is what the
    # reference's algorithm gives (ids are identical), rising with ef.
    d = ((Q[:200, None, :].astype(np.float64) - X[None, :, :]) ** 2).sum(-1)
    truth = np.argsort(d, axis=1)[:, :10]
    r256 = workloads.recall(ix.search_rows(Q[:200], 10, 256)["label"], truth, 10)
    r512 = workloads.recall(ix.search_rows(Q[:200], 10, 512)["label"], truth, 10)
    assert 0.85 <= r256 <= r512


def test_edge_cases(lib):
    empty = lib.Index(8)
    out = empty.search_rows(np.zeros((2, 8), np.float32), 3, 10)
    assert (out["label"] == lib.INVALID_LABEL).all() and np.isinf(out["distance"]).all()
    assert len(empty.search_one(np.zeros(8, np.float32), 5)) == 0

    rng = np.random.default_rng(1)
    five = rng.uniform(-1, 1, (5, 4)).astype(np.float32)
    ix = lib.Index(4)
    ix.add(five)
    res = ix.search_one(five[0], 50, 1)  # ef < k is clamped to k
    assert len(res) == 5 and res["label"][0] == 0
    assert np.all(np.diff(res["distance"]) >= 0)
    with pytest.raises(ValueError, match="k must be > 0"):
        ix.search_one(five[0], 0)

    same = np.full((300, 12), 0.5, np.float32)
    ix = lib.Index(12)
    ix.add(same, num_threads=1)
    res = ix.search_one(same[0], 10, 64)
    assert len(res) == 10 and (res["distance"] == 0).all() and len(set(res["label"].tolist())) == 10

    ix = lib.Index(4)
    ix.add(rng.uniform(-1, 1, (20, 4)).astype(np.float32))
    for lab in range(20):
        ix.mark_deleted(lab)
    assert len(ix.search_one(np.zeros(4, np.float32), 5)) == 0

    # Every node reachable: exhaustive ef returns all n.
    line = np.full((100, 16), 0.5, np.float32)
    line[:, 0] = np.arange(100)
    ix = lib.Index(16)
    ix.add(line, num_threads=4)
    assert len(ix.search_one(line[0], 100, 200)) == 100
    res = ix.search_one(line[0], 10, 64, filter=np.array([7, 99]))
    assert sorted(res["label"].tolist()) == [7, 99]


def test_line_known_answer(lib):
    line = np.arange(200, dtype=np.float32).reshape(-1, 1)
    ix = lib.Index(1)
    for i in range(200):
        ix.add(line[i:i + 1], [i])
    assert ix.search_one(np.array([42.2], np.float32), 3, 50)["label"].tolist() == [42, 43, 41]


def test_soft_delete_and_filter_semantics(lib):
    rng = np.random.default_rng(55)
    centers = rng.standard_normal((32, 16)).astype(np.float32) * 4
    X = (centers[rng.integers(0, 32, 3000)] + rng.standard_normal((3000, 16))).astype(np.float32)
    ix = lib.Index(16)
    ix.add(X)
    deleted = set(range(0, 3000, 3))
    for lab in deleted:
        ix.mark_deleted(lab)
    for i in range(0, 600, 3):
        res = ix.search_one(X[i], 10, 64)
        assert len(res) == 10 and not (set(res["label"].tolist()) & deleted)
    allow = np.zeros(3000, bool)
    allow[::2] = True
    res = ix.search_one(X[0], 20, 100, filter=allow)
    assert all(allow[l] and l not in deleted for l in res["label"].tolist())


def test_device_buffers_match_host_api(golden, lib):
    ix = lib.Index.load(GOLDEN / "l2_d200_avx512.idx")
    Q = golden["l2_d200/queries"]
    want = ix.search_rows(Q, 10, 64)
    dev = ix.device()
    q = lib.DeviceBuffer.of(Q)
    out = lib.DeviceBuffer(Q.shape[0] * 10 * lib.NEIGHBOR.itemsize)
    stream = lib.Stream()
    dev.search_device(q.ptr, Q.shape[0], 10, 64, out.ptr, stream=stream.handle)
    stream.synchronize()
    got = out.download(np.empty(Q.shape[0] * 10, lib.NEIGHBOR)).reshape(Q.shape[0], 10)
    assert rows_equal(got, want)


def test_work_counters_equal_oracle_counters(golden, lib, oracle):
    """The kernel's own per-query work counts (which give the roofline bytes)
    equal the oracle's counts of the same traversal."""
    path = GOLDEN / "l2_d32_scalar.idx"
    ix, orc = lib.Index.load(path), oracle.load(path)
    Q = golden["l2_d32/queries"]
    dev = ix.device()
    dev.set_option("collect_stats", 1)
    ix.search_rows(Q, 10, 64)
    got = dev.read_stats(Q.shape[0]).astype(np.uint64)
    dev.set_option("collect_stats", 0)
    _, want = orc.search(Q, 10, 64, order="avx512", counters=True)
    assert np.array_equal(got, want)


def test_deletions_side_list_kernel(golden, lib, oracle):
    """Soft deletions alone run the unfiltered traversal with a side list of
    admitted deleted nodes ("deleted_path" 2 forces it at any deletion count,
    0 uses the filtered kernel): both equal the reference's golden rows and
    the oracle, and a side list that overflows hands its queries on exactly."""
    for name in case_names(golden):
        print("case", name, flush=True)
        ix = lib.Index.load(GOLDEN / f"{name}_avx512.idx")
        Q = golden[f"{name}/queries"]
        for lab in golden[f"{name}/deleted"].tolist():
            ix.mark_deleted(lab)
        dev = ix.device()
        for path in (2, 0, 1):
            dev.set_option("deleted_path", path)
            assert rows_equal(ix.search_rows(Q, 10, 64), golden[f"{name}/avx512/deleted/k10_ef64"]), (name, path)
        path = GOLDEN / f"{name}_scalar.idx"
        ix, orc = lib.Index.load(path), oracle.load(path)
        for lab in golden[f"{name}/deleted"].tolist():
            ix.mark_deleted(lab)
            orc.mark_deleted(lab)
        ix.device().set_option("deleted_path", 2)
        for k, ef in ((1, 1), (10, 10), (10, 64), (7, 300)):
            assert rows_equal(ix.search_rows(Q, k, ef), orc.search(Q, k, ef, order="avx512")), (name, k, ef)

    path = GOLDEN / "l2_d32_scalar.idx"
    ix, orc = lib.Index.load(path), oracle.load(path)
    Q = golden["l2_d32/queries"]
    for lab in range(0, len(ix), 2):
        ix.mark_deleted(lab)
        orc.mark_deleted(lab)
    dev = ix.device()
    dev.set_option("deleted_path", 2)
    for ef in (64, 200, 500):
        rows = ix.search_rows(Q, 10, ef)
        handed = dev.get_option("last_handoff_candidates")
        print("half deleted, ef", ef, "handed on", handed, flush=True)
        assert rows_equal(rows, orc.search(Q, 10, ef, order="avx512")), ef
    assert handed > 0

    # Work counts and a reduced copy: the side-list kernel and the filtered
    # kernel run the same traversal.
    ix = lib.Index.load(GOLDEN / "l2_d32_scalar.idx")
    for lab in golden["l2_d32/deleted"].tolist():
        ix.mark_deleted(lab)
    dev = ix.device()
    dev.set_option("collect_stats", 1)
    got = {}
    for st in (0, 1):
        dev.set_option("storage", st)
        for path in (2, 0):
            dev.set_option("deleted_path", path)
            got[st, path] = (ix.search_rows(Q, 10, 100), dev.read_stats(Q.shape[0]).copy())
        assert rows_equal(got[st, 2][0], got[st, 0][0]), st
        assert np.array_equal(got[st, 2][1], got[st, 0][1]), st
    dev.set_option("collect_stats", 0)
