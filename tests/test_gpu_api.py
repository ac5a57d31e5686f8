"""The boundary around search_batch on the GPU (through the C ABI).

* Filters other than a BitsetFilter (include/hnsw/filter.hpp:12-30, a
  FunctionFilter or any subclass; eligible(id) = !deleted[id] &&
  allowed(labels[id]), src/index_impl.hpp:259-261): a predicate given to the
  C++ API (hnsw_index_search_batch_pred) and an id bitmap given to the device
  layer (hnswgpu_search_batch_ids) both return the oracle's rows for the
  equivalent label bitset, with deletions.
* Concurrent search_batch calls on one index from several host threads (the
  reference's searches are concurrent, include/hnsw/index.hpp:31-36,
  tests/cpp/test_concurrency.cpp:75-98) return the sequential rows.
* An index loaded with mmap=True (src/persist.cpp:477-503) searches like the
  golden rows.
* After an add, the device copy is refreshed in place: handles given out
  before stay valid, keep their options, and search the new graph exactly.
* Device-buffer cosine searches prepare queries with the integer emulation of
  the reference's double arithmetic and still return the golden rows.
* A host-buffer search that fails after queueing its copies leaves its call
  context usable: larger searches afterwards return the oracle's rows.
"""

from __future__ import annotations

import threading

import numpy as np
import pytest

from conftest import GOLDEN, case_names

pytestmark = pytest.mark.gpu


def rows_equal(got, want):
    return np.array_equal(got["label"], want["label"]) and np.array_equal(
        got["distance"].view(np.uint32), want["distance"].view(np.uint32))


def label_words(allowed_labels, size):
    w = np.zeros(max(1, (size + 63) // 64), np.uint64)
    for lab in allowed_labels:
        w[lab >> 6] |= np.uint64(1) << np.uint64(lab & 63)
    return w


def test_predicate_filter_equals_label_bitset(golden, lib, oracle):
    def pred(label):
        return label % 7 == 3

    for name in case_names(golden):
        print("case", name, flush=True)
        path = GOLDEN / f"{name}_avx512.idx"
        ix, orc = lib.Index.load(path), oracle.load(path)
        Q = golden[f"{name}/queries"]
        labels = golden[f"{name}/labels"].astype(np.int64)
        size = int(labels.max()) + 1
        words = (label_words([int(l) for l in labels if pred(int(l))], size), size)
        for deleted in (False, True):
            if deleted:
                for lab in golden[f"{name}/deleted"].tolist():
                    ix.mark_deleted(lab)
                    orc.mark_deleted(lab)
            for k, ef in ((1, 10), (10, 64), (25, 200)):
                want = orc.search(Q, k, ef, order="avx512", filter=words)
                got = ix.search_rows(Q, k, ef, filter=pred)
                assert rows_equal(got, want), (name, deleted, k, ef)


def test_predicate_filter_exceptions_propagate(golden, lib):
    ix = lib.Index.load(GOLDEN / "l2_d32_avx512.idx")
    Q = golden["l2_d32/queries"]

    def bad(label):
        raise KeyError(label)

    with pytest.raises(KeyError):
        ix.search_rows(Q, 10, 64, filter=bad)


def test_id_bitmap_entry_on_20k_with_deletions(lib, oracle, tmp_path):
    import workloads

    X = workloads.uniform(20000, 128, 0)
    Q = workloads.uniform(400, 128, 1)
    rng = np.random.default_rng(3)
    perm = rng.permutation(20000).astype(np.uint64) * 3 + 11  # labels unlike internal ids
    ix = lib.Index(128, "l2")
    ix.add(X, ids=perm, num_threads=0)
    for lab in perm[::97].tolist():
        ix.mark_deleted(lab)
    ix.save(tmp_path / "p.idx")
    orc = oracle.load(tmp_path / "p.idx")
    allowed = (perm % 7 == 3) | (perm % 5 == 0)  # by internal id: node i has label perm[i]
    size = int(perm.max()) + 1
    words = (label_words(perm[allowed].astype(np.int64).tolist(), size), size)
    dev = ix.device()
    for k, ef in ((10, 64), (10, 256), (100, 400)):
        want = orc.search(Q, k, ef, order="avx512", filter=words)
        rows, counts = dev.search_rows_ids(Q, k, ef, id_mask=allowed)
        assert rows_equal(rows, want), (k, ef)
        assert np.array_equal(counts, (want["label"] != lib.INVALID_LABEL).sum(1))
        # The C++ API's predicate path takes the same route.
        got = ix.search_rows(Q, k, ef, filter=lambda l: (l % 7 == 3) or (l % 5 == 0))
        assert rows_equal(got, want), (k, ef)
    # No bitmap: deletions only.
    rows, _ = dev.search_rows_ids(Q, 10, 64, id_mask=None)
    assert rows_equal(rows, orc.search(Q, 10, 64, order="avx512"))


def test_concurrent_search_batch_on_one_index(lib):
    import workloads

    X = workloads.mixture_128(20000, seed=21)
    Q = workloads.mixture_128(2000, seed=1021)
    ix = lib.Index(128, "l2")
    ix.add(X, num_threads=0)
    rng = np.random.default_rng(4)
    allow = rng.random(20000) < 0.2
    jobs = [(Q[i * 250:(i + 1) * 250], int(k), int(ef), f)
            for i, (k, ef, f) in enumerate([(10, 64, None), (1, 16, None), (100, 200, allow), (10, 128, None),
                                            (10, 32, allow), (50, 512, None), (10, 64, allow), (5, 300, None)])]
    want = [ix.search_rows(q, k, ef, filter=f) for q, k, ef, f in jobs]
    got = [[None] * len(jobs) for _ in range(4)]
    errors = []

    def worker(t):
        try:
            for rep in range(3):
                for j in range(len(jobs)):
                    jj = (j + t + rep) % len(jobs)
                    q, k, ef, f = jobs[jj]
                    got[t][jj] = ix.search_rows(q, k, ef, filter=f)
                    assert rows_equal(got[t][jj], want[jj]), (t, jj)
        except BaseException as e:  # reported below
            errors.append(e)

    ts = [threading.Thread(target=worker, args=(t,)) for t in range(4)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors[0]
    dev = ix.device()
    assert 1 <= dev.get_option("contexts") <= dev.get_option("max_contexts")


def test_device_searches_on_two_streams_overlap_safely(lib):
    """Asynchronous device-buffer searches alternating over two streams take
    two contexts of the handle (their launches may overlap on the device) and
    each returns the rows of the same search alone, with deletions and, for
    some, a label filter; calls on one stream keep reusing one context."""
    import workloads

    X = workloads.mixture_128(20000, seed=25)
    Q = workloads.mixture_128(3000, seed=1025)
    ix = lib.Index(128, "l2")
    ix.add(X, num_threads=0)
    for label in range(3, 20000, 101):
        ix.mark_deleted(label)
    dev = ix.device()
    efs = [16, 64, 128, 256, 100, 40]
    # Every third search with a 30 % label filter (eligibility bitmaps in the context's scratch).
    fw, fsize = lib.filter_words(np.random.default_rng(6).random(20000) < 0.3)
    d_f = lib.DeviceBuffer.of(fw)
    filt = [i % 3 == 2 for i in range(len(efs))]
    want = [ix.search_rows(Q, 10, ef, filter=(fw, fsize) if f else None) for ef, f in zip(efs, filt)]
    before = dev.get_option("contexts")
    d_q = lib.DeviceBuffer.of(np.ascontiguousarray(Q, np.float32))
    streams = [lib.Stream(), lib.Stream()]
    outs = [lib.DeviceBuffer(Q.shape[0] * 10 * lib.NEIGHBOR.itemsize) for _ in efs]
    for rep in range(2):
        for i, ef in enumerate(efs):
            dev.search_device(d_q.ptr, Q.shape[0], 10, ef, outs[i].ptr, stream=streams[i % 2].handle,
                              d_filter=d_f.ptr if filt[i] else 0, filter_size=fsize if filt[i] else 0)
        for s in streams:
            s.synchronize()
        for i in range(len(efs)):
            rows = outs[i].download(np.empty(Q.shape[0] * 10, lib.NEIGHBOR)).reshape(Q.shape[0], 10)
            assert rows_equal(rows, want[i]), (rep, efs[i])
    assert before <= dev.get_option("contexts") <= min(before + 2, dev.get_option("max_contexts"))
    n = dev.get_option("contexts")
    one = lib.Stream()
    for ef in efs:
        dev.search_device(d_q.ptr, Q.shape[0], 10, ef, outs[0].ptr, stream=one.handle)
    one.synchronize()
    assert dev.get_option("contexts") <= n + 1


def test_mmap_loaded_index_equals_golden_rows(golden, lib):
    for name in ("l2_d32", "cos_d100", "ties_d8"):
        ix = lib.Index.load(GOLDEN / f"{name}_avx512.idx", mmap=True)
        assert ix.read_only
        Q = golden[f"{name}/queries"]
        filt = (golden[f"{name}/allow_words"], int(golden[f"{name}/allow_size"][0]))
        for key in golden.files:
            if key.startswith(f"{name}/avx512/plain/") or key.startswith(f"{name}/avx512/filter/"):
                k, ef = [int(v[1:] if v[0] == "k" else v[2:]) for v in key.split("/")[-1].split("_")]
                got = ix.search_rows(Q, k, ef, filter=filt if "/filter/" in key else None)
                assert rows_equal(got, golden[key]), key


def test_add_refreshes_the_device_copy_in_place(lib, oracle, tmp_path):
    import workloads

    X = workloads.mixture_128(12000, seed=23)
    Q = workloads.mixture_128(300, seed=1023)
    ix = lib.Index(128, "l2")
    ix.add(X[:8000], num_threads=0)
    dev = ix.device()
    dev.set_option("storage", 1)
    dev.set_option("waves_per_cu", 12)
    before = ix.search_rows(Q, 10, 64)
    ix.add(X[8000:], num_threads=0)
    after = ix.search_rows(Q, 10, 128)
    # The handle taken before the add is still the index's, with its options.
    assert dev.get_option("storage") == 1 and dev.get_option("waves_per_cu") == 12
    rows, _ = dev.search_rows(Q, 10, 128)
    assert rows_equal(rows, after)
    assert (after["label"] >= 8000).any() and not (before["label"] >= 8000).any()
    dev.set_option("storage", 0)
    ix.save(tmp_path / "a.idx")
    orc = oracle.load(tmp_path / "a.idx")
    assert rows_equal(ix.search_rows(Q, 10, 128), orc.search(Q, 10, 128, order="avx512"))


def test_device_buffer_cosine_equals_golden_rows(golden, lib):
    """Device-buffer searches compute the cosine scale on the device without
    FP64; the rows equal the reference's (AVX-512 build) bit for bit."""
    ix = lib.Index.load(GOLDEN / "cos_d100_avx512.idx")
    Q = np.ascontiguousarray(golden["cos_d100/queries"], np.float32)
    dev = ix.device()
    stream = lib.Stream()
    d_q = lib.DeviceBuffer.of(Q)
    for key in golden.files:
        if not key.startswith("cos_d100/avx512/plain/"):
            continue
        k, ef = [int(v[1:] if v[0] == "k" else v[2:]) for v in key.split("/")[-1].split("_")]
        d_out = lib.DeviceBuffer(Q.shape[0] * k * lib.NEIGHBOR.itemsize)
        dev.search_device(d_q.ptr, Q.shape[0], k, ef, d_out.ptr, stream=stream.handle)
        stream.synchronize()
        got = d_out.download(np.empty(Q.shape[0] * k, lib.NEIGHBOR)).reshape(Q.shape[0], k)
        assert rows_equal(got, golden[key]), key
    # Mixed launches (device queries) as well.
    res = dev.search_mixed(Q, [10] * Q.shape[0], [64] * Q.shape[0])
    want = golden["cos_d100/avx512/plain/k10_ef64"]
    for i, (rows, n) in enumerate(res):
        assert rows_equal(rows, want[i]), i


def test_failed_call_then_larger_calls_stay_exact(lib, oracle, tmp_path):
    """Host-buffer searches stage small batches through pinned buffers of the
    call's context. A call that fails after queueing its copies (ef beyond 32
    bits) must leave that context usable: the following calls, whose batches
    grow the staging buffers, return the oracle's rows."""
    import workloads

    X = workloads.mixture_128(8000, seed=71)
    Q = workloads.mixture_128(1200, seed=1071)
    ix = lib.Index(128, "l2")
    ix.add(X, num_threads=0)
    path = tmp_path / "m8k.idx"
    ix.save(path)
    orc = oracle.load(path)
    dev = ix.device()
    for nq in (1, 300, 1200):
        print("case nq", nq, flush=True)
        with pytest.raises(ValueError, match="out of range"):
            dev.search_rows(Q[:nq], 10, 1 << 33)
        got, counts = dev.search_rows(Q[:nq], 10, 64)
        assert rows_equal(got, orc.search(Q[:nq], 10, 64, order="avx512")), nq
        assert (counts == 10).all()
