"""GPU parity on the benchmark configurations at their own shapes.

* C2: 200,000 x 128 integer Gaussian mixture (exact distance ties are common),
  10,000 queries, k 10, ef 16..256, unfiltered fp32 search_batch: every row
  equals the oracle in the AVX-512 order (ids, order, distance bits). The same
  with 1 % soft-deleted (side-list and filtered kernels) and a 10 % filter.
* C3: the 1,000-centre cosine mixture at 100,000 x 100: fp32 equals the
  oracle at k 10 / 100 and ef 32..512; the bf16 traversal returns the exact
  fp32 distance of every id it returns (oracle, bit for bit), sorted by
  (distance, id), with recall@k within 0.01 of the fp32 traversal's.
* C5: the serving mix on the 200,000 mixture index: requests of 1..512
  queries (log-uniform), k {1, 10, 100}, ef {32..512}, 30 % with a 10 %-density
  filter, 1 % of labels soft-deleted, fp32 and bf16 mixed; every request
  equals itself alone, and the fp32 requests equal the oracle.

Algorithm: src/search_layer.cpp:29-194 of the reference (greedy descent,
Alg. 2, batch padding); filters src/index_impl.hpp:259-261.
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rows_equal(got, want):
    return np.array_equal(got["label"], want["label"]) and np.array_equal(
        got["distance"].view(np.uint32), want["distance"].view(np.uint32))


def mismatching_rows(got, want):
    bad = ~((got["label"] == want["label"]).all(1) &
            (got["distance"].view(np.uint32) == want["distance"].view(np.uint32)).all(1))
    return int(bad.sum())


@pytest.fixture(scope="module")
def c2(lib, oracle, tmp_path_factory):
    import workloads

    X = workloads.mixture_128(200_000, seed=2)
    Q = workloads.mixture_128(10_000, seed=1002)
    ix = lib.Index(128, "l2", M=16, ef_construction=200)
    ix.add(X, num_threads=0)
    path = tmp_path_factory.mktemp("c2") / "c2.idx"
    ix.save(path)
    return ix, oracle.load(path), X, Q, path


def test_c2_mixture_200k_unfiltered_equals_oracle(c2):
    ix, orc, X, Q, _ = c2
    for ef in (16, 32, 64, 128, 256):
        got = ix.search_rows(Q, 10, ef)
        want = orc.search(Q, 10, ef, order="avx512")
        print("ef", ef, "mismatching rows", mismatching_rows(got, want), flush=True)
        assert rows_equal(got, want), ef


def test_c2_mixture_200k_deleted_and_filtered_equal_oracle(lib, oracle, c2):
    _, _, X, Q, path = c2
    ix, orc = lib.Index.load(path), oracle.load(path)
    rng = np.random.default_rng(52)
    for lab in rng.choice(X.shape[0], X.shape[0] // 100, replace=False).tolist():
        ix.mark_deleted(lab)
        orc.mark_deleted(lab)
    dev = ix.device()
    Qs = Q[:4000]
    for ef in (16, 64, 128, 256):
        want = orc.search(Qs, 10, ef, order="avx512")
        for path_opt in (1, 0):  # side list of deleted nodes; the filtered kernel
            dev.set_option("deleted_path", path_opt)
            got = ix.search_rows(Qs, 10, ef)
            assert rows_equal(got, want), (ef, path_opt, mismatching_rows(got, want))
    dev.set_option("deleted_path", 1)
    allow = rng.random(X.shape[0]) < 0.10
    w = np.zeros((X.shape[0] + 63) // 64, np.uint64)
    idx = np.nonzero(allow)[0].astype(np.uint64)
    np.bitwise_or.at(w, (idx >> np.uint64(6)).astype(np.int64), np.uint64(1) << (idx & np.uint64(63)))
    filt = (w, X.shape[0])
    for ef in (64, 256):
        want = orc.search(Qs[:2000], 10, ef, order="avx512", filter=filt)
        got = ix.search_rows(Qs[:2000], 10, ef, filter=filt)
        assert rows_equal(got, want), (ef, mismatching_rows(got, want))


def test_c3_cosine_100k_fp32_equals_oracle_and_bf16_is_exact(lib, oracle, tmp_path):
    import workloads

    X = workloads.cosine_100(100_000, seed=3)
    Q = workloads.cosine_100(2_000, seed=1003)
    ix = lib.Index(100, "cosine", M=16, ef_construction=200)
    ix.add(X, num_threads=0)
    ix.save(tmp_path / "c3.idx")
    orc = oracle.load(tmp_path / "c3.idx")
    bf = lib.BruteForceIndex(100, "cosine")
    bf.add(X)
    truth = bf.search_rows(Q, 100)["label"].astype(np.int64)
    del bf
    dev = ix.device()
    for k in (10, 100):
        for ef in (32, 64, 128, 256, 512):
            if ef < k:
                continue
            dev.set_option("storage", 0)
            exact = ix.search_rows(Q, k, ef)
            want = orc.search(Q, k, ef, order="avx512")
            assert rows_equal(exact, want), (k, ef, mismatching_rows(exact, want))
            dev.set_option("storage", 1)
            low = ix.search_rows(Q, k, ef)
            lab = low["label"]
            n_valid = (lab != lib.INVALID_LABEL).sum(1)
            assert np.all(n_valid == k), (k, ef)
            d = orc.label_dists(Q, lab, order="avx512")
            assert np.array_equal(low["distance"].view(np.uint32), d.view(np.uint32)), (k, ef)
            key = low["distance"].astype(np.float64) * 2.0 ** 32 + lab.astype(np.float64)
            assert np.all(np.diff(low["distance"].astype(np.float64), axis=1) >= 0), (k, ef)
            del key
            r32 = workloads.recall(exact["label"].astype(np.int64), truth, k)
            r16 = workloads.recall(lab.astype(np.int64), truth, k)
            print(f"k={k} ef={ef} recall fp32 {r32:.4f} bf16 {r16:.4f}", flush=True)
            assert r16 >= r32 - 0.01, (k, ef, r32, r16)
    dev.set_option("storage", 0)


def test_c5_serving_mix_on_200k(lib, oracle, c2):
    _, _, X, Q, path = c2
    ix, orc = lib.Index.load(path), oracle.load(path)
    rng = np.random.default_rng(55)
    for lab in rng.choice(X.shape[0], X.shape[0] // 100, replace=False).tolist():
        ix.mark_deleted(lab)
        orc.mark_deleted(lab)
    allow = rng.random(X.shape[0]) < 0.10
    w = np.zeros((X.shape[0] + 63) // 64, np.uint64)
    idx = np.nonzero(allow)[0].astype(np.uint64)
    np.bitwise_or.at(w, (idx >> np.uint64(6)).astype(np.int64), np.uint64(1) << (idx & np.uint64(63)))
    filt = (w, X.shape[0])
    dev = ix.device()
    srv = lib.Server(dev, slots=4, max_batch=2048, storages=(1,))
    reqs = []
    for _ in range(60):
        nq = int(np.exp(rng.uniform(0.0, np.log(512.0))))
        s = int(rng.integers(0, Q.shape[0] - nq))
        reqs.append((Q[s:s + nq], int(rng.choice([1, 10, 100])), int(rng.choice([32, 64, 128, 256, 512])),
                     filt if rng.random() < 0.3 else None, 1 if rng.random() < 0.5 else 0))
    try:
        tickets = [srv.submit(q, k, ef, f, storage=st) for q, k, ef, f, st in reqs]
        got = [srv.wait(t) for t in tickets]
    finally:
        srv.close()
    n_fp32 = 0
    for i, ((q, k, ef, f, st), (rows, counts)) in enumerate(zip(reqs, got)):
        dev.set_option("storage", st)
        alone, alone_counts = dev.search_rows(q, k, ef, f)
        dev.set_option("storage", 0)
        assert rows_equal(rows, alone), i
        assert np.array_equal(counts, alone_counts), i
        if st == 0:
            n_fp32 += 1
            want = orc.search(q, k, ef, order="avx512", filter=f)
            assert rows_equal(rows, want), (i, k, ef, f is not None)
    assert n_fp32 > 10
