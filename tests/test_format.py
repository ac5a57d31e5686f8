"""Host side of the product on CPU: the reference file format and the builder.

* the builder (1 thread) writes files byte-identical to the reference's own
  (tests/golden/*_avx512.idx were written by the reference under its AVX-512
  kernels, the order the engine computes in);
* load (read or mmap) + save round-trips byte for byte;
* every truncation and single-bit flip is rejected with the reference's error
  text (tests/cpp/test_persist.cpp:110-162 of the reference);
* argument errors keep the reference's exception classes and messages.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import GOLDEN, case_names


def test_builder_writes_reference_bytes(golden, lib, tmp_path):
    for name in case_names(golden):
        X, lab = golden[f"{name}/data"], golden[f"{name}/labels"]
        metric, M, efc = (int(v) for v in golden[f"{name}/meta"])
        ix = lib.Index(X.shape[1], lib.METRIC_NAMES[metric], M=M, ef_construction=efc, seed=100)
        ix.add(X, lab, num_threads=1)
        out = tmp_path / f"{name}.idx"
        ix.save(out)
        assert out.read_bytes() == (GOLDEN / f"{name}_avx512.idx").read_bytes(), name


@pytest.mark.parametrize("mmap", [False, True])
def test_load_save_round_trip(golden, lib, tmp_path, mmap):
    for name in case_names(golden):
        src = GOLDEN / f"{name}_scalar.idx"
        ix = lib.Index.load(src, mmap=mmap)
        assert len(ix) == golden[f"{name}/data"].shape[0]
        assert ix.read_only == mmap
        out = tmp_path / f"{name}.idx"
        ix.save(out)
        assert out.read_bytes() == src.read_bytes()


def test_parallel_build_loads_and_is_complete(golden, lib, tmp_path):
    X = golden["l2_d32/data"]
    ix = lib.Index(32, "l2", M=16, ef_construction=100)
    ix.add(X, num_threads=4)
    ix.save(tmp_path / "p.idx")
    back = lib.Index.load(tmp_path / "p.idx")
    assert len(back) == X.shape[0]
    for lab in (0, 17, X.shape[0] - 1):
        assert lab in back
        np.testing.assert_array_equal(back.get_vector(lab), X[lab])


def test_truncation_and_bit_flips_are_rejected(lib, tmp_path):
    data = (GOLDEN / "ip_d16_avx512.idx").read_bytes()
    bad = tmp_path / "bad.idx"
    rng = np.random.default_rng(5)
    cuts = sorted(set([0, 1, 8, 100, 255, 256, 300, len(data) // 2, len(data) - 1] +
                      rng.integers(1, len(data), 20).tolist()))
    for cut in cuts:
        bad.write_bytes(data[:cut])
        for mmap in (False, True):
            with pytest.raises(RuntimeError):
                lib.Index.load(bad, mmap=mmap)
    positions = list(range(256)) + rng.integers(256, len(data), 200).tolist()
    for pos in positions:
        b = bytearray(data)
        b[pos] ^= 1 << int(rng.integers(0, 8))
        bad.write_bytes(bytes(b))
        with pytest.raises(RuntimeError, match="corrupt or incompatible index file"):
            lib.Index.load(bad, mmap=bool(pos & 1))


def test_parallel_build_is_deterministic_and_keeps_recall(lib, oracle, tmp_path):
    """Multi-threaded builds insert in waves against the graph frozen at each
    wave's start: the file is the same for any thread count >= 2, and recall
    at a given ef stays within 0.01 of the one-thread (reference-order) build."""
    import workloads

    X = workloads.mixture_128(8000, seed=11, dim=32)
    Q = workloads.mixture_128(300, seed=1011, dim=32)
    files = {}
    for threads in (1, 2, 5, 8):
        ix = lib.Index(32, "l2", M=8, ef_construction=64)
        ix.add(X, num_threads=threads)
        files[threads] = tmp_path / f"t{threads}.idx"
        ix.save(files[threads])
    assert files[2].read_bytes() == files[5].read_bytes() == files[8].read_bytes()
    assert files[2].read_bytes() != files[1].read_bytes()
    truth = oracle.brute_force(X, Q, 10)["label"].astype(np.int64)
    for ef in (10, 20, 40):
        r1 = workloads.recall(oracle.load(files[1]).search(Q, 10, ef)["label"].astype(np.int64), truth, 10)
        r8 = workloads.recall(oracle.load(files[8]).search(Q, 10, ef)["label"].astype(np.int64), truth, 10)
        assert r8 >= r1 - 0.01, (ef, r1, r8)


def test_reference_loads_our_files(golden, lib, reference, tmp_path):
    X = golden["cos_d100/data"]
    ix = lib.Index(100, "cosine", M=16, ef_construction=100)
    ix.add(X, num_threads=3)
    ix.mark_deleted(5)
    ix.save(tmp_path / "ours.idx")
    r = reference.load(str(tmp_path / "ours.idx"), mmap=True)
    r.save(str(tmp_path / "back.idx"))
    assert (tmp_path / "back.idx").read_bytes() == (tmp_path / "ours.idx").read_bytes()


def test_argument_errors_match_reference(lib):
    with pytest.raises(ValueError, match="dim must be > 0"):
        lib.Index(0)
    with pytest.raises(ValueError, match="M must be >= 2"):
        lib.Index(4, M=1)
    with pytest.raises(ValueError, match="ef_construction must be > 0"):
        lib.Index(4, ef_construction=0)
    ix = lib.Index(4)
    with pytest.raises(ValueError, match="set_ef: ef must be > 0"):
        ix.set_ef(0)
    v = np.ones((2, 4), np.float32)
    ix.add(v[:1], [7])
    with pytest.raises(ValueError, match="already exists"):
        ix.add(v[:1], [7])
    with pytest.raises(ValueError, match="duplicate label 9 in batch"):
        ix.add(v, [9, 9])
    assert len(ix) == 1 and 9 not in ix
    with pytest.raises(ValueError, match="unknown label"):
        ix.mark_deleted(12345)
    with pytest.raises(ValueError, match="search: k must be > 0"):
        ix.search_one(v[0], 0)
    fixed = lib.Index(4, max_elements=2)
    fixed.add(v)
    with pytest.raises(RuntimeError, match="Index is full"):
        fixed.add(v[:1], [99])


def test_mmap_index_is_read_only(lib):
    ix = lib.Index.load(GOLDEN / "l2_d32_avx512.idx", mmap=True)
    with pytest.raises(RuntimeError, match="read-only"):
        ix.add(np.zeros((1, 32), np.float32), [123456])
    ix.mark_deleted(3)  # tombstones live in memory
    assert ix.is_deleted(3)
    ix.set_ef(77)
    assert ix.ef == 77
