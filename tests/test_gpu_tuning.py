"""The autotuner: every class it is given gets a measured choice, and search
results do not depend on the launch shape (ids, distance bits)."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rows_equal(a, b):
    return np.array_equal(a["label"], b["label"]) and np.array_equal(a["distance"].view(np.uint32),
                                                                      b["distance"].view(np.uint32))


def test_autotune_choices_and_exactness(lib):
    import workloads

    X = workloads.mixture_128(20_000, seed=51)
    Q = workloads.mixture_128(600, seed=1051)
    ix = lib.Index(128, "l2")
    ix.add(X)
    dev = ix.device()
    before = {ef: ix.search_rows(Q, 10, ef) for ef in (32, 100, 200, 400)}
    filt = np.random.default_rng(2).random(X.shape[0]) < 0.2
    before_f = ix.search_rows(Q, 10, 150, filter=filt)
    report = dev.autotune(Q, 10, [32, 100, 200, 400])
    assert report.count("best ef=") == 4
    for cls in (64, 128, 256, 512):
        assert dev.get_option(f"tuned_waves_per_cu_{cls}") in (8, 12, 16, 20, 24)
        assert dev.get_option(f"tuned_hash_log2_{cls}") in (10, 11, 12, 13)
        assert dev.get_option(f"tuned_us_{cls}") > 0
    assert dev.get_option("use_tuned") == 1
    for ef, want in before.items():
        assert rows_equal(ix.search_rows(Q, 10, ef), want), ef
    assert rows_equal(ix.search_rows(Q, 10, 150, filter=filt), before_f)
    # Every shape of the sweep gives the same rows.
    for wpc, hl, gather in ((8, 13, 1), (24, 11, 0)):
        dev.set_option("use_tuned", 0)
        dev.set_option("waves_per_cu", wpc)
        dev.set_option("hash_log2", hl)
        dev.set_option("generic_gather", gather)
        assert rows_equal(ix.search_rows(Q, 10, 100), before[100])
    dev.set_option("hash_log2", 0)
    dev.set_option("generic_gather", 0)
    dev.set_option("waves_per_cu", 16)
    with pytest.raises(ValueError):
        dev.autotune(Q[:0], 10, [64])
