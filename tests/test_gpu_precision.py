"""Reduced-precision traversal (bf16, fp8 e4m3 + per-vector scale) with exact
fp32 re-rank.

* the stored bf16 rows are the round-to-nearest-even bf16 of the fp32 rows;
* the stored fp8 rows are the OCP e4m3 encoding (bias 7, largest finite 448,
  round to nearest even) of x * (448 / max|x|), written here from the format's
  definition, with scale = max|x| / 448;
* every returned distance is the exact fp32 distance of its id (the oracle's,
  bit for bit), rows are sorted by (distance, id), and recall stays within a
  stated margin of the fp32 traversal. Distance error of the stored copies
  against FP64 is measured and bounded.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def bf16_rne(x: np.ndarray) -> np.ndarray:
    u = x.astype(np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def e4m3_table():
    vals = np.zeros(256, np.float64)
    for code in range(256):
        s, e, m = code >> 7, (code >> 3) & 0xF, code & 7
        if e == 0xF and m == 7:
            vals[code] = np.nan
            continue
        v = (m / 8.0) * 2.0 ** -6 if e == 0 else (1 + m / 8.0) * 2.0 ** (e - 7)
        vals[code] = -v if s else v
    return vals


def e4m3_encode(x: np.ndarray) -> np.ndarray:
    table = e4m3_table()
    pos = np.array([c for c in range(128) if not np.isnan(table[c])])  # codes 0..126
    pv = table[pos]
    a = np.abs(x.astype(np.float64))
    idx = np.clip(np.searchsorted(pv, a), 1, len(pv) - 1)
    lo, hi = pos[idx - 1], pos[idx]
    dlo, dhi = a - table[lo], table[hi] - a
    pick = np.where(dlo < dhi, lo, np.where(dhi < dlo, hi, np.where(lo % 2 == 0, lo, hi)))
    pick = np.where(a >= pv[-1], pos[-1], pick)
    return (pick | np.where(np.signbit(x), 128, 0)).astype(np.uint8)


def test_bf16_rows_are_round_to_nearest_even(golden, lib):
    ix = lib.Index.load(GOLDEN / "l2_d200_avx512.idx")
    dev = ix.device()
    dev.set_option("storage", 1)
    for node in range(0, 300, 7):
        row32, _, _ = dev.get_row(node, 0, 208)
        row16, _, norm = dev.get_row(node, 1, 208)
        assert np.array_equal(row16, bf16_rne(row32))
        back = (row16.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
        assert abs(norm - (back ** 2).sum()) <= 1e-5 * max(1.0, (back ** 2).sum())


def test_fp8_rows_are_ocp_e4m3_with_scale(golden, lib):
    ix = lib.Index.load(GOLDEN / "l2_d200_avx512.idx")
    dev = ix.device()
    dev.set_option("storage", 2)
    table = e4m3_table()
    for node in range(0, 300, 7):
        row32, _, _ = dev.get_row(node, 0, 208)
        row8, scale, norm = dev.get_row(node, 2, 208)
        m = np.float32(np.abs(row32).max())
        assert scale == np.float32(m / np.float32(448.0))
        scaled = np.clip(row32 * (np.float32(448.0) / m), -448, 448).astype(np.float32)
        assert np.array_equal(row8, e4m3_encode(scaled)), node
        deq = table[row8] * scale
        assert abs(norm - (deq ** 2).sum()) <= 1e-5 * (deq ** 2).sum()
        # Stored-value error against FP64: half an e4m3 step (2^-4 relative) of the scale range.
        assert np.all(np.abs(deq - row32.astype(np.float64)) <= 2.0 ** -4 * np.abs(row32) + scale * 2.0 ** -9 + 1e-30)


@pytest.mark.parametrize("storage", [1, 2])
def test_reduced_traversal_returns_exact_distances(golden, lib, oracle, storage):
    """Dimension 128 (the fixed row width): ef 64 (insert re-rank) and ef 400
    (the largest capacity class, re-ranked in place and sorted) return the
    exact fp32 distances of their ids, sorted by (distance, id), at the recall
    of the fp32 traversal."""
    import workloads

    X = workloads.mixture_128(20000, seed=9)
    Q = workloads.mixture_128(300, seed=1009)
    ix = lib.Index(128, "l2")
    ix.add(X)
    dev = ix.device()
    # Exact in float64: the data are integers below 256.
    q64, x64 = Q.astype(np.float64), X.astype(np.float64)
    d64 = (q64 ** 2).sum(1)[:, None] + (x64 ** 2).sum(1)[None, :] - 2.0 * q64 @ x64.T
    for k, ef in ((10, 64), (10, 400), (100, 400)):
        exact = ix.search_rows(Q, k, ef)
        dev.set_option("storage", storage)
        got = ix.search_rows(Q, k, ef)
        dev.set_option("storage", 0)
        lab = got["label"].astype(np.int64)
        # Distances are the exact fp32 distances (AVX-512 order) of the returned ids.
        want = np.array([[oracle.l2(Q[i], X[j]) for j in lab[i]] for i in range(Q.shape[0])], np.float32)
        assert np.array_equal(got["distance"].view(np.uint32), want.view(np.uint32)), (k, ef)
        key = got["distance"].astype(np.float64) * 2 ** 32 + lab
        assert np.all(np.diff(key, axis=1) > 0), (k, ef)
        truth = np.argsort(d64, axis=1, kind="stable")[:, :k]
        r_exact = workloads.recall(exact["label"], truth, k)
        r_low = workloads.recall(got["label"], truth, k)
        assert r_low >= r_exact - 0.02, (k, ef, r_low, r_exact)


def test_reduced_cosine_and_filters(golden, lib):
    ix = lib.Index.load(GOLDEN / "cos_d100_avx512.idx")
    Q = golden["cos_d100/queries"]
    filt = (golden["cos_d100/allow_words"], int(golden["cos_d100/allow_size"][0]))
    exact = ix.search_rows(Q, 10, 200, filter=filt)
    dev = ix.device()
    for storage in (1, 2):
        dev.set_option("storage", storage)
        got = ix.search_rows(Q, 10, 200, filter=filt)
        # ef = 200 on 800 nodes finds the same answer set; distances are exact fp32.
        assert np.array_equal(got["label"], exact["label"]), storage
        assert np.array_equal(got["distance"].view(np.uint32), exact["distance"].view(np.uint32))
    dev.set_option("storage", 0)


def test_high_dim_960_fp32_parity_and_fp8_rerank(lib, oracle, tmp_path):
    """Padded dim 960 (the generic row loop, 60 chunks in groups of 8): the fp32
    traversal equals the oracle bit for bit, and the fp8 e4m3 traversal returns
    the exact fp32 distances of its ids with recall close to the fp32 one."""
    import workloads

    X = workloads.sparse_960(3000, seed=14)
    Q = workloads.sparse_960(120, seed=1014)
    ix = lib.Index(960, "l2", M=16, ef_construction=100)
    ix.add(X, num_threads=0)
    ix.save(tmp_path / "s960.idx")
    orc = oracle.load(tmp_path / "s960.idx")
    for ef in (10, 64, 200):
        got = ix.search_rows(Q, 10, ef)
        want = orc.search(Q, 10, ef, order="avx512")
        assert np.array_equal(got["label"], want["label"]), ef
        assert np.array_equal(got["distance"].view(np.uint32), want["distance"].view(np.uint32)), ef
    exact = ix.search_rows(Q, 10, 64)
    dev = ix.device()
    dev.set_option("storage", 2)
    low = ix.search_rows(Q, 10, 64)
    dev.set_option("storage", 0)
    lab = low["label"].astype(np.int64)
    want = np.array([[oracle.l2(Q[i], X[j]) for j in lab[i]] for i in range(Q.shape[0])], np.float32)
    assert np.array_equal(low["distance"].view(np.uint32), want.view(np.uint32))
    d64 = ((Q[:, None, :].astype(np.float64) - X[None, :, :]) ** 2).sum(-1)
    truth = np.argsort(d64, axis=1, kind="stable")[:, :10]
    assert workloads.recall(low["label"], truth, 10) >= workloads.recall(exact["label"], truth, 10) - 0.03


def test_reduced_largest_class_generic_width(golden, lib):
    """ef above 256 (the largest capacity class) at a dimension of no fixed
    row width (100, padded to 112): plain, filtered and deletions-only
    searches on the bf16 and fp8 copies return exact fp32 distances, sorted,
    and the answer set of the fp32 search on 800 nodes (k 10 and 100)."""
    ix = lib.Index.load(GOLDEN / "cos_d100_avx512.idx")
    Q = golden["cos_d100/queries"]
    filt = (golden["cos_d100/allow_words"], int(golden["cos_d100/allow_size"][0]))
    dev = ix.device()
    for deleted in (False, True):
        if deleted:
            for lab in golden["cos_d100/deleted"].tolist():
                ix.mark_deleted(lab)
        for f in (None, filt):
            for k in (10, 100):
                dev.set_option("storage", 0)
                exact = ix.search_rows(Q, k, 400, filter=f)
                for storage in (1, 2):
                    dev.set_option("storage", storage)
                    got = ix.search_rows(Q, k, 400, filter=f)
                    case = (deleted, f is not None, k, storage)
                    print("case", case, flush=True)
                    hit = 0
                    for i in range(Q.shape[0]):
                        want = dict(zip(exact["label"][i].tolist(), exact["distance"][i].view(np.uint32).tolist()))
                        n = int((got["label"][i] != lib.INVALID_LABEL).sum())
                        assert np.all(got["label"][i][n:] == lib.INVALID_LABEL), case
                        assert np.all(np.diff(got["distance"][i][:n].astype(np.float64)) >= 0), case
                        for lab, d in zip(got["label"][i][:n].tolist(), got["distance"][i][:n].view(np.uint32).tolist()):
                            if lab in want:
                                hit += 1
                                assert d == want[lab], case
                    valid = int((exact["label"] != lib.INVALID_LABEL).sum())
                    assert hit >= 0.99 * valid, (case, hit, valid)
    dev.set_option("storage", 0)
