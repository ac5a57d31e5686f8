"""Filtered search at low selectivity: the candidate set outgrows its
registers and continues in the per-wavefront HBM overflow (sorted runs), with
no handoff to the general kernel, and the rows stay equal to the oracle's
(AVX-512 summation order, exact Alg. 2 order)."""

from __future__ import annotations

import os
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rows_equal(a, b):
    return np.array_equal(a["label"].astype(np.uint64), b["label"].astype(np.uint64)) and np.array_equal(
        a["distance"].view(np.uint32), b["distance"].view(np.uint32))


@pytest.fixture(scope="module")
def built(lib, oracle, tmp_path_factory):
    import workloads

    X = workloads.mixture_128(20_000, seed=41)
    Q = workloads.mixture_128(120, seed=1041)
    ix = lib.Index(128, "l2")
    ix.add(X)
    path = tmp_path_factory.mktemp("filt") / "f.idx"
    ix.save(path)
    return lib.Index.load(path), oracle.load(path), X, Q


@pytest.mark.parametrize("density", [0.10, 0.05, 0.02])
def test_low_selectivity_equals_oracle_without_handoff(lib, built, density):
    ix, orc, X, Q = built
    rng = np.random.default_rng(int(density * 1000))
    w, size = lib.filter_words(rng.random(X.shape[0]) < density)
    dev = ix.device()
    for ef, k in ((16, 10), (64, 10), (128, 100), (256, 10), (512, 10)):
        got = ix.search_rows(Q, k, ef, filter=(w, size))
        handoffs = dev.get_option("last_handoff_candidates")
        want = orc.search(Q, k, ef, order="avx512", filter=(w, size))
        assert rows_equal(got, want), (density, ef, k)
        if density >= 0.05:
            assert handoffs == 0, (density, ef, handoffs)


def test_overflow_and_deletions_together(lib, built):
    ix, orc, X, Q = built
    rng = np.random.default_rng(9)
    w, size = lib.filter_words(rng.random(X.shape[0]) < 0.05)
    # The general kernel (HBM heaps, no register capacity) is the reference for the same index state.
    dev = ix.device()
    for label in range(0, X.shape[0], 50):
        ix.mark_deleted(label)
    fast = ix.search_rows(Q, 10, 200, filter=(w, size))
    dev.set_option("force_spill", 1)
    slow = ix.search_rows(Q, 10, 200, filter=(w, size))
    dev.set_option("force_spill", 0)
    assert rows_equal(fast, slow)


def test_dense_label_filter_takes_the_side_list(lib, oracle, built):
    """A label filter allowing almost every label (here 97 %, with 1 % of the
    nodes soft-deleted on top) runs the unfiltered traversal with a side list
    of the admitted ineligible nodes (the host estimates the eligible fraction
    from the filter words): rows equal the oracle's, and equal the filtered
    kernel's (option deleted_path 0)."""
    ix, orc, X, Q = built
    rng = np.random.default_rng(97)
    allow = rng.random(X.shape[0]) < 0.97
    w, size = lib.filter_words(allow)
    # A copy of the module's index, so that its deletions stay local.
    tmp = tempfile.mkdtemp()
    ix.save(os.path.join(tmp, "d.idx"))
    ix2, orc2 = lib.Index.load(os.path.join(tmp, "d.idx")), oracle.load(os.path.join(tmp, "d.idx"))
    for lab in rng.choice(X.shape[0], X.shape[0] // 100, replace=False).tolist():
        ix2.mark_deleted(lab)
        orc2.mark_deleted(lab)
    dev = ix2.device()
    for ef, k in ((16, 10), (64, 10), (128, 100), (256, 10), (512, 10)):
        want = orc2.search(Q, k, ef, order="avx512", filter=(w, size))
        dev.set_option("deleted_path", 1)
        got = ix2.search_rows(Q, k, ef, filter=(w, size))
        assert rows_equal(got, want), (ef, k)
        dev.set_option("deleted_path", 0)
        assert rows_equal(ix2.search_rows(Q, k, ef, filter=(w, size)), want), (ef, k)
    dev.set_option("deleted_path", 1)
