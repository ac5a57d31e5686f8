"""Writes the golden vectors under tests/golden/ from the reference itself.

The reference library is compiled from its own sources into oracle/_ref
(make -C oracle ref); this script drives it through oracle/ref_shim.cpp:
  * builds small indexes with the reference's Index::add_batch (1 thread, so
    the graph is deterministic) under its scalar and AVX-512 kernels and saves
    them with Index::save -> <case>_<isa>.idx (reference-format files);
  * runs the reference's Index::search_batch on them for several k / ef /
    filter / deletion settings and BruteForceIndex on the raw data;
  * stores inputs and expected (label, distance) rows in golden.npz.

Run from the repo root:  python tests/golden/make_golden.py
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "oracle"))
from oracle import Reference  # noqa: E402

OUT = Path(__file__).resolve().parent


def clustered(n, dim, seed, clusters=16, spread=4.0):
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((clusters, dim)).astype(np.float32) * np.float32(spread)
    return (c[rng.integers(0, clusters, n)] + rng.standard_normal((n, dim)).astype(np.float32)).astype(np.float32)


def words_of(allowed_labels, size):
    w = np.zeros(max(1, (size + 63) // 64), dtype=np.uint64)
    for lab in allowed_labels:
        w[lab >> 6] |= np.uint64(1) << np.uint64(lab & 63)
    return w


def cases():
    rng = np.random.default_rng(7)
    ints = rng.integers(0, 4, (1500, 8)).astype(np.float32)  # many exact distance ties
    line = np.arange(200, dtype=np.float32).reshape(-1, 1)
    return [
        # name, data, queries, metric, M, efc, labels
        ("l2_d32", clustered(2000, 32, 11), clustered(40, 32, 12), 0, 16, 100, None),
        ("cos_d100", clustered(800, 100, 21), clustered(30, 100, 22), 2, 16, 100, None),
        ("ip_d16", clustered(1200, 16, 31), clustered(30, 16, 32), 1, 8, 80, None),
        ("ties_d8", ints, rng.integers(0, 4, (40, 8)).astype(np.float32), 0, 12, 60, None),
        ("line_d1", line, np.array([[42.2], [0.0], [199.5], [77.7]], dtype=np.float32), 0, 16, 200, None),
        ("l2_d3_labels", clustered(800, 3, 41), clustered(20, 3, 42), 0, 16, 100,
         (np.arange(800, dtype=np.uint64) * 7 + 1000)),
        ("l2_d200", clustered(300, 200, 51), clustered(16, 200, 52), 0, 16, 100, None),
    ]


def main() -> None:
    ref = Reference()
    store = {}
    for name, X, Q, metric, M, efc, labels in cases():
        n = X.shape[0]
        lab = np.arange(n, dtype=np.uint64) if labels is None else labels
        store[f"{name}/data"] = X
        store[f"{name}/queries"] = Q
        store[f"{name}/labels"] = lab
        store[f"{name}/meta"] = np.array([metric, M, efc], dtype=np.int64)
        # Deleted labels (every 13th) and a 10 % allow-list by label.
        deleted = lab[::13].copy()
        allow = lab[(np.arange(n) * 2654435761 % 1000) < 100]
        fsize = int(lab.max()) + 1
        store[f"{name}/deleted"] = deleted
        store[f"{name}/allow_words"] = words_of(allow.tolist(), fsize)
        store[f"{name}/allow_size"] = np.array([fsize], dtype=np.int64)
        for isa in ("scalar", "avx512"):
            ref.set_isa(isa)
            path = OUT / f"{name}_{isa}.idx"
            ref.build_save(str(path), X, labels=lab, metric=metric, M=M, efc=efc, seed=100, threads=1)
            idx = ref.load(str(path))
            for k, ef in ((1, 10), (10, 10), (10, 64), (25, 200)):
                if k > n:
                    continue
                out = idx.search(Q, k, ef, threads=1)
                store[f"{name}/{isa}/plain/k{k}_ef{ef}"] = out
                out = idx.search(Q, k, ef, threads=1, filter=(store[f"{name}/allow_words"], fsize))
                store[f"{name}/{isa}/filter/k{k}_ef{ef}"] = out
            didx = ref.load(str(path))
            for d in deleted.tolist():
                didx.mark_deleted(d)
            for k, ef in ((10, 64),):
                store[f"{name}/{isa}/deleted/k{k}_ef{ef}"] = didx.search(Q, k, ef, threads=1)
                store[f"{name}/{isa}/deleted_filter/k{k}_ef{ef}"] = didx.search(
                    Q, k, ef, threads=1, filter=(store[f"{name}/allow_words"], fsize))
            store[f"{name}/{isa}/brute/k10"] = ref.brute_force(X, Q, min(10, n), metric=metric, labels=lab)
            store[f"{name}/{isa}/brute_filter/k10"] = ref.brute_force(
                X, Q, min(10, n), metric=metric, labels=lab, filter=(store[f"{name}/allow_words"], fsize))
    np.savez_compressed(OUT / "golden.npz", **store)
    total = sum(p.stat().st_size for p in OUT.glob("*.idx")) + (OUT / "golden.npz").stat().st_size
    print(f"wrote {len(store)} arrays, {total / 1e6:.2f} MB")


if __name__ == "__main__":
    main()
