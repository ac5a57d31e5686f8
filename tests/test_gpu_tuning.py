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
    for i, cls in enumerate((64, 128, 256, 512)):
        assert dev.get_option(f"tuned_waves_per_cu_{cls}") in (8, 12, 16, 20, 24)
        assert dev.get_option(f"tuned_hash_log2_{cls}") in (10, 11, 12, 13)
        assert dev.get_option(f"tuned_row_trip_{cls}") in (1, 2)
        assert dev.get_option(f"tuned_min_class_{cls}") in (i, min(i + 1, 3))
        assert dev.get_option(f"tuned_us_{cls}") > 0
    # The report covers both stage-2 dimensions: each class measured both row trips and a larger class.
    for ef in (32, 100, 200):
        lines = [l for l in report.splitlines() if l.startswith(f"ef={ef} ")]
        assert {l.split("row_trip=")[1].split()[0] for l in lines} == {"1", "2"}, ef
        assert len({l.split("min_class=")[1].split()[0] for l in lines}) == 2, ef
    assert dev.get_option("use_tuned") == 1
    for ef, want in before.items():
        assert rows_equal(ix.search_rows(Q, 10, ef), want), ef
    assert rows_equal(ix.search_rows(Q, 10, 150, filter=filt), before_f)
    # Every shape of the sweep gives the same rows: occupancy, hash size, row loop, rows per
    # memory round trip and capacity class, unfiltered and filtered, fp32, bf16 and fp8.
    dev.set_option("use_tuned", 0)
    for wpc, hl, gather, trip, cls in ((8, 13, 1, 1, 0), (24, 11, 0, 2, 0), (16, 11, 0, 1, 2), (16, 12, 0, 2, 3)):
        dev.set_option("waves_per_cu", wpc)
        dev.set_option("hash_log2", hl)
        dev.set_option("generic_gather", gather)
        dev.set_option("row_trip", trip)
        dev.set_option("min_class", cls)
        assert rows_equal(ix.search_rows(Q, 10, 100), before[100]), (wpc, hl, gather, trip, cls)
        assert rows_equal(ix.search_rows(Q, 10, 32), before[32]), (wpc, hl, gather, trip, cls)
        assert rows_equal(ix.search_rows(Q, 10, 150, filter=filt), before_f), (wpc, hl, gather, trip, cls)
    dev.set_option("min_class", 0)
    for storage in (1, 2):
        dev.set_option("storage", storage)
        dev.set_option("row_trip", 0)
        want = ix.search_rows(Q, 10, 100)
        for trip in (1, 2):
            dev.set_option("row_trip", trip)
            assert rows_equal(ix.search_rows(Q, 10, 100), want), (storage, trip)
    dev.set_option("storage", 0)
    dev.set_option("row_trip", 0)
    for bad in (("row_trip", 3), ("min_class", 4)):
        with pytest.raises(ValueError):
            dev.set_option(*bad)
    dev.set_option("hash_log2", 0)
    dev.set_option("generic_gather", 0)
    dev.set_option("waves_per_cu", 16)
    with pytest.raises(ValueError):
        dev.autotune(Q[:0], 10, [64])
