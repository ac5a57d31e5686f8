"""Graph construction with the insertion searches on the GPU
(hnsw_index_add_batch_device), against the host builder.

The builder inserts in waves against the graph frozen at each wave's start
(src/index.cpp:205-306 of the reference is the one-node-per-wave case). A
wave's phase 1 (greedy descent + the ef_construction beam, src/search_layer.cpp:
29-134) runs on the GPU for its level-0 nodes once waves reach 512 nodes; the
GPU computes the host's summation order and Alg. 2's order and rules, so the
index file must be byte-identical to the host build with the same thread count
(which the host builder's own tests pin to the reference's bytes for one
thread and to recall for several).
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def build_pair(lib, X, metric, tmp_path, M=16, efc=200, threads=8, split=None):
    files = []
    stats = None
    for device in (False, True):
        ix = lib.Index(X.shape[1], metric, M=M, ef_construction=efc)
        parts = [X] if split is None else [X[:split], X[split:]]
        first = 0
        for part in parts:
            labels = np.arange(first, first + part.shape[0], dtype=np.uint64)
            st = ix.add(part, labels, num_threads=threads, device=device)
            first += part.shape[0]
            if device:
                stats = st
        path = tmp_path / f"{metric}_{device}.idx"
        ix.save(path)
        files.append(path)
    return files[0].read_bytes(), files[1].read_bytes(), stats


@pytest.mark.parametrize("metric,dim", [("l2", 32), ("ip", 24), ("cosine", 100)])
def test_device_build_equals_host_build(lib, tmp_path, metric, dim):
    import workloads

    if metric == "cosine":
        X = workloads.cosine_100(30_000, seed=21)
    else:
        X = workloads.mixture_128(30_000, seed=21, dim=dim)
        if metric == "ip":
            X = X / np.float32(255.0)
    host, device, st = build_pair(lib, X, metric, tmp_path)
    print(metric, st, flush=True)
    assert st["waves"] > 0 and st["device_nodes"] > 5_000, st
    assert host == device, metric


def test_device_build_of_a_second_batch_and_small_m(lib, tmp_path):
    """A second add onto a built graph (the device copy starts from a graph
    with upper layers and back-links), M = 6, and many integer ties."""
    import workloads

    X = workloads.mixture_128(36_000, seed=22, dim=16, n_centers=32, sigma=4.0)
    host, device, st = build_pair(lib, X, "l2", tmp_path, M=6, efc=64, threads=4, split=20_000)
    print(st, flush=True)
    assert st["device_nodes"] > 5_000, st
    assert host == device


def test_device_build_beyond_the_fast_kernel_ef(lib, tmp_path):
    """ef_construction above 512: the beams run in the general kernel."""
    import workloads

    X = workloads.mixture_128(20_000, seed=23, dim=16)
    host, device, st = build_pair(lib, X, "l2", tmp_path, M=8, efc=520, threads=8)
    print(st, flush=True)
    assert st["device_nodes"] > 1_000, st
    assert host == device


def test_device_built_index_searches_like_the_host_built_one(lib, oracle, tmp_path):
    """Recall of a device build against FP64 ground truth, and its search rows
    against the oracle on the same file."""
    import workloads

    X = workloads.mixture_128(50_000, seed=24)
    Q = workloads.mixture_128(500, seed=1024)
    ix = lib.Index(128, "l2", M=16, ef_construction=200)
    ix.add(X, num_threads=8, device=True)
    path = tmp_path / "dev.idx"
    ix.save(path)
    truth = oracle.brute_force(X, Q, 10)["label"].astype(np.int64)
    got = ix.search_rows(Q, 10, 64)
    want = oracle.load(path).search(Q, 10, 64, order="avx512")
    assert np.array_equal(got["label"], want["label"])
    assert workloads.recall(got["label"].astype(np.int64), truth, 10) >= 0.95
