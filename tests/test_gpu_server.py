"""Serving loop (hnswgpu_server_*): requests of mixed batch size, k, ef,
filter and storage submitted together are batched into shared launches, and
each result equals the same request run alone through search_batch (ids,
distance bits, counts). Graph replays equal eager launches."""

from __future__ import annotations

import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rows_equal(a, b):
    return np.array_equal(a["label"], b["label"]) and np.array_equal(a["distance"].view(np.uint32),
                                                                      b["distance"].view(np.uint32))


@pytest.fixture(scope="module")
def served(lib):
    import workloads

    X = workloads.mixture_128(20_000, seed=31)
    Q = workloads.mixture_128(3_000, seed=1031)
    ix = lib.Index(128, "l2")
    ix.add(X)
    dev = ix.device()
    rng = np.random.default_rng(7)
    filters = [rng.random(X.shape[0]) < 0.1, rng.random(X.shape[0]) < 0.5]
    return ix, dev, X, Q, filters


def make_requests(rng, Q, n, filters, storages=(0, 1)):
    reqs = []
    for _ in range(n):
        nq = int(np.exp(rng.uniform(0, np.log(300))))
        start = int(rng.integers(0, Q.shape[0] - nq))
        fid = int(rng.integers(-1, len(filters)))
        reqs.append(dict(q=Q[start:start + nq], k=int(rng.choice([1, 10, 100])),
                         ef=int(rng.choice([0, 32, 64, 128, 300, 600])), fid=fid,
                         storage=int(rng.choice(storages))))
    return reqs


def alone(dev, r, filters):
    dev.set_option("storage", r["storage"])
    rows, counts = dev.search_rows(r["q"], r["k"], r["ef"], None if r["fid"] < 0 else filters[r["fid"]])
    dev.set_option("storage", 0)
    return rows, counts


def run_server(srv, reqs, filters, threads=1):
    tickets = [None] * len(reqs)

    def submit(idx):
        for i in idx:
            r = reqs[i]
            tickets[i] = srv.submit(r["q"], r["k"], r["ef"], None if r["fid"] < 0 else filters[r["fid"]],
                                    storage=r["storage"])

    parts = [list(range(t, len(reqs), threads)) for t in range(threads)]
    ts = [threading.Thread(target=submit, args=(p,)) for p in parts]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    return [srv.wait(t) for t in tickets]


def test_each_result_same_alone_and_in_the_mix(lib, served):
    ix, dev, X, Q, filters = served
    rng = np.random.default_rng(11)
    reqs = make_requests(rng, Q, 80, filters)
    srv = lib.Server(dev, slots=3, max_batch=1024, storages=(1,))
    got = run_server(srv, reqs, filters, threads=4)
    st = srv.stats()
    srv.close()
    assert st["requests"] == len(reqs) and st["queries"] == sum(r["q"].shape[0] for r in reqs)
    assert st["launches"] < len(reqs)  # requests shared launches
    assert 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    for r, (rows, counts) in zip(reqs, got):
        want, wc = alone(dev, r, filters)
        assert rows_equal(rows, want), (r["k"], r["ef"], r["fid"], r["storage"])
        assert np.array_equal(counts, wc)


def test_wait_latency_and_heavy_classes(lib, served):
    """Per-request latency readout, and heavy requests (expensive queries:
    filtered at high ef) in launches of their own: results unchanged, and the
    latencies are the ones the statistics count."""
    ix, dev, X, Q, filters = served
    rng = np.random.default_rng(12)
    reqs = make_requests(rng, Q, 40, filters)
    srv = lib.Server(dev, slots=3, max_batch=1024, storages=(1,), heavy_cost=256, max_cost=1 << 16)
    try:
        tickets = [srv.submit(r["q"], r["k"], r["ef"], None if r["fid"] < 0 else filters[r["fid"]],
                              storage=r["storage"]) for r in reqs]
        outs = [srv._live[t][2] for t in tickets]
        lat = [srv.wait_latency(t) for t in tickets]
        st = srv.stats()
    finally:
        srv.close()
    assert all(l > 0 for l in lat)
    assert abs(max(lat) - st["max_ms"]) < 1e-6 and st["requests"] == len(reqs)
    for r, out in zip(reqs, outs):
        want, _ = alone(dev, r, filters)
        assert rows_equal(out.reshape(r["q"].shape[0], -1), want), (r["k"], r["ef"], r["fid"], r["storage"])


def test_graph_replays_equal_eager_launches(lib, served):
    ix, dev, X, Q, filters = served
    rng = np.random.default_rng(12)
    reqs = make_requests(rng, Q, 30, filters, storages=(0,))
    eager = lib.Server(dev, slots=2, max_batch=512, graphs=False)
    want = run_server(eager, reqs, filters)
    eager.close()
    srv = lib.Server(dev, slots=2, max_batch=512, graphs=True)
    for round_ in range(3):
        got = run_server(srv, reqs, filters)
        for (a, ac), (b, bc) in zip(got, want):
            assert rows_equal(a, b) and np.array_equal(ac, bc), round_
        # Eager work of the program between replays.
        rows, _ = dev.search_rows(Q[:50], 10, 64)
        assert rows.shape == (50, 10)
    assert srv.stats()["graph_launches"] > 0
    srv.close()


def test_large_requests_are_split(lib, served):
    ix, dev, X, Q, filters = served
    srv = lib.Server(dev, slots=2, max_batch=64, max_rows=64 * 10)
    t1 = srv.submit(Q[:500], 10, 64, filters[1])
    t2 = srv.submit(Q[500:520], 100, 128)  # 100 rows per query: 6 queries per launch at most
    rows1, c1 = srv.wait(t1)
    rows2, c2 = srv.wait(t2)
    assert srv.stats()["launches"] >= 500 // 64 + 20 // 6
    srv.close()
    want1, w1 = dev.search_rows(Q[:500], 10, 64, filters[1])
    want2, w2 = dev.search_rows(Q[500:520], 100, 128)
    assert rows_equal(rows1, want1) and np.array_equal(c1, w1)
    assert rows_equal(rows2, want2) and np.array_equal(c2, w2)


def test_deletions_and_errors(lib):
    import workloads

    X = workloads.mixture_128(5_000, seed=32)
    Q = workloads.mixture_128(200, seed=1032)
    ix = lib.Index(128, "l2")
    ix.add(X)
    for label in range(0, 5_000, 13):
        ix.mark_deleted(label)
    dev = ix.device()
    srv = lib.Server(dev, slots=2, max_batch=256, max_filter_labels=5_000)
    mask = np.random.default_rng(3).random(5_000) < 0.4
    # The filtered kernel's rows; then served through it (deleted_path 0) and
    # through the side-list kernel for deletions alone (2).
    dev.set_option("deleted_path", 0)
    wants = [dev.search_rows(Q[i * 20:(i + 1) * 20], 10, 50, mask if i % 2 else None) for i in range(10)]
    for path in (0, 2):
        dev.set_option("deleted_path", path)
        ts = [srv.submit(Q[i * 20:(i + 1) * 20], 10, 50, mask if i % 2 else None) for i in range(10)]
        for i, t in enumerate(ts):
            rows, counts = srv.wait(t)
            want, wc = wants[i]
            assert rows_equal(rows, want) and np.array_equal(counts, wc), (path, i)
            assert not np.isin(rows["label"][rows["label"] != lib.INVALID_LABEL].astype(np.int64) % 13, [0]).any()
    dev.set_option("deleted_path", 1)
    with pytest.raises(ValueError):
        srv.submit(Q[:2], 0, 10)
    with pytest.raises(ValueError):
        srv.submit(Q[:2], 10, 10, storage=2)  # fp8 copy not built
    with pytest.raises(ValueError):
        srv.submit(Q[:2], 10, 10, np.ones(10_000, bool))  # more labels than the server was sized for
    srv.close()


def test_deletions_after_graphs_were_captured(lib):
    """A live server with graphs: unfiltered requests before any deletion,
    after a few (the side-list kernel, a graph of its own) and after many
    (past 1 in 16: the filtered kernel) equal search_batch at each stage and
    never return a deleted label."""
    import workloads

    X = workloads.mixture_128(6_000, seed=33)
    Q = workloads.mixture_128(300, seed=1033)
    ix = lib.Index(128, "l2")
    ix.add(X)
    dev = ix.device()
    srv = lib.Server(dev, slots=2, max_batch=512, graphs=True)
    deleted = set()
    for stage, step in enumerate((0, 97, 7)):
        if step:
            for label in range(stage, 6_000, step):
                ix.mark_deleted(label)
                deleted.add(label)
        print("stage", stage, "deleted", len(deleted), flush=True)
        for replay in range(2):
            ts = [srv.submit(Q[i * 50:(i + 1) * 50], 10, ef) for i, ef in enumerate((32, 64, 128, 200, 300, 500))]
            for i, t in enumerate(ts):
                rows, counts = srv.wait(t)
                want, wc = dev.search_rows(Q[i * 50:(i + 1) * 50], 10, (32, 64, 128, 200, 300, 500)[i])
                assert rows_equal(rows, want) and np.array_equal(counts, wc), (stage, replay, i)
                assert not (set(rows["label"].ravel().tolist()) & deleted), (stage, replay, i)
    assert srv.stats()["graph_launches"] > 0
    srv.close()


def test_graphs_follow_option_and_tuning_changes(lib, served):
    """Options that change the launch shape or reallocate scratch (waves per CU,
    visited-table and hash sizes, the autotuner's shapes) between server
    launches: a graph captured before is captured again instead of replaying
    stale scratch pointers, and every result still equals the request alone."""
    ix, dev, X, Q, filters = served
    rng = np.random.default_rng(13)
    srv = lib.Server(dev, slots=2, max_batch=512, storages=(1,))
    settings = [{}, {"waves_per_cu": 12}, {"visit_log2": 14}, {"hash_log2": 12}, {"wide_hash": 1},
                {"hash_log2": 0, "wide_hash": 0, "waves_per_cu": 20, "visit_log2": 15}]
    try:
        for step, opts in enumerate(settings + ["autotune"]):
            if opts == "autotune":
                dev.autotune(Q[:512], 10, [64, 128, 256, 512])
            else:
                for name, value in opts.items():
                    dev.set_option(name, value)
            print("step", step, opts, flush=True)
            reqs = make_requests(rng, Q, 24, filters)
            got = run_server(srv, reqs, filters)
            for i, (r, (rows, counts)) in enumerate(zip(reqs, got)):
                want_rows, want_counts = alone(dev, r, filters)
                assert rows_equal(rows, want_rows), (step, i)
                assert np.array_equal(counts, want_counts), (step, i)
        st = srv.stats()
        assert st["recaptures"] > 0 and st["graph_launches"] > 0
        # The index cannot be changed under a server.
        with pytest.raises(ValueError, match="server is attached"):
            ix.add(X[:10])
    finally:
        srv.close()
        for name, value in (("use_tuned", 0), ("waves_per_cu", 16), ("hash_log2", 0), ("visit_log2", 15),
                            ("wide_hash", 0)):
            dev.set_option(name, value)
