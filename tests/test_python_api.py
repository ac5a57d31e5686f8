"""The Python module's argument rules, which need no GPU: the behaviours the
reference's binding tests (tests/py/test_index.py:109-191, :228-233,
:235-252, :273-291) check before any search runs, on this module.

* input conversion: non-numeric queries are a TypeError, wrong shapes a
  ValueError naming the shape, k = 0 a ValueError;
* constructor arguments (dim 0, unknown metric, M < 2) are ValueErrors;
* labels: default arange continuing from the current size, custom integer
  ids, negative / wrong-length / non-integer / duplicate / existing ids
  rejected with the reference's exception types;
* get_vector returns the stored (cosine-normalised) row and leaves the
  caller's buffer alone;
* filters of the wrong kind are TypeErrors;
* mark_deleted of unknown or negative labels is a ValueError;
* an mmap-loaded index refuses add; corrupt, truncated and missing files are
  RuntimeErrors;
* an empty index answers with -1 / inf padding.
"""

from __future__ import annotations

import numpy as np
import pytest


def clustered(n: int, dim: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    centres = rng.normal(0.0, 3.0, size=(12, dim))
    return (centres[rng.integers(0, 12, size=n)] + rng.normal(size=(n, dim))).astype(np.float32)


@pytest.fixture(scope="module")
def small(lib):
    idx = lib.Index(dim=32, metric="l2")
    idx.add(clustered(500, 32, 11), num_threads=2)
    return idx


def test_query_conversion_errors(lib, small):
    with pytest.raises(ValueError, match="shape"):
        small.search(np.zeros((3, 31), np.float32))
    with pytest.raises(ValueError):
        small.search(np.zeros(31, np.float32))
    with pytest.raises(ValueError):
        small.search(np.zeros((2, 2, 32), np.float32))
    with pytest.raises(ValueError, match="k must be"):
        small.search(np.zeros(32, np.float32), k=0)
    with pytest.raises(TypeError):
        small.search(np.array(["a", "b"]))
    idx = lib.Index(dim=8)
    with pytest.raises(ValueError):
        idx.add(np.zeros((4, 7), np.float32))


def test_constructor_errors(lib):
    with pytest.raises(ValueError):
        lib.Index(dim=0)
    with pytest.raises(ValueError):
        lib.Index(dim=8, metric="hamming")
    with pytest.raises(ValueError):
        lib.Index(dim=8, M=1)


def test_labels_default_and_custom(lib):
    base = clustered(300, 8, 5)
    idx = lib.Index(dim=8)
    idx.add(base[:100], ids=np.arange(1000, 1100))
    idx.add(base[100:200])  # continues at len(idx): labels 100..199
    assert 1000 in idx and 150 in idx and 99 not in idx and 1100 not in idx
    idx.add(base[200:202], ids=np.array([7, 8], dtype=np.uint64))
    assert 7 in idx and 8 in idx and len(idx) == 202
    np.testing.assert_array_equal(idx.get_vector(1005), base[5])


def test_label_errors(lib):
    idx = lib.Index(dim=4)
    v = np.random.default_rng(0).random((3, 4), dtype=np.float32)
    with pytest.raises(ValueError, match="non-negative"):
        idx.add(v, ids=[1, -2, 3])
    with pytest.raises(ValueError, match="shape"):
        idx.add(v, ids=[1, 2])
    with pytest.raises(TypeError):
        idx.add(v, ids=[1.5, 2.5, 3.5])
    with pytest.raises(ValueError, match="duplicate"):
        idx.add(v, ids=[1, 1, 2])
    idx.add(v, ids=[1, 2, 3])
    with pytest.raises(ValueError, match="already exists"):
        idx.add(v[:1], ids=[2])
    assert len(idx) == 3


def test_get_vector_is_the_normalised_row(lib):
    v = np.array([[3.0, 4.0]], np.float32)
    idx = lib.Index(dim=2, metric="cosine")
    idx.add(v)
    np.testing.assert_allclose(idx.get_vector(0), [0.6, 0.8], rtol=1e-6)
    np.testing.assert_array_equal(v, [[3.0, 4.0]])


def test_filter_kind_errors(lib, small):
    q = clustered(2, 32, 12)
    with pytest.raises(TypeError):
        small.search(q, k=3, filter=np.array([0.5, 1.5]))
    with pytest.raises(TypeError):
        small.search(q, k=3, filter="nope")
    with pytest.raises(ValueError):
        small.search(q, k=3, filter=np.zeros((2, 2), bool))
    with pytest.raises(ValueError, match="non-negative"):
        small.search(q, k=3, filter=[3, -1])


def test_mark_deleted_errors(lib, small):
    with pytest.raises(ValueError):
        small.mark_deleted(5000)
    with pytest.raises(ValueError):
        small.mark_deleted(-1)
    assert not small.is_deleted(3)


def test_read_only_and_bad_files(lib, small, tmp_path):
    path = tmp_path / "index.bin"
    small.save(str(path))
    ro = lib.Index.load(str(path), mmap=True)
    assert ro.read_only and len(ro) == len(small)
    with pytest.raises(RuntimeError, match="read-only"):
        ro.add(clustered(1, 32, 13), ids=[999999])
    raw = bytearray(path.read_bytes())
    raw[len(raw) // 2] ^= 0xFF
    bad = tmp_path / "bad.bin"
    bad.write_bytes(bytes(raw))
    for mmap in (False, True):
        with pytest.raises(RuntimeError, match="corrupt"):
            lib.Index.load(str(bad), mmap=mmap)
    (tmp_path / "short.bin").write_bytes(path.read_bytes()[:1000])
    with pytest.raises(RuntimeError):
        lib.Index.load(str(tmp_path / "short.bin"))
    with pytest.raises(RuntimeError):
        lib.Index.load(str(tmp_path / "missing.bin"))


def test_empty_index_pads(lib):
    idx = lib.Index(dim=3)
    ids, dists = idx.search(np.zeros(3, np.float32), k=2)
    assert ids.shape == (2,) and (ids == -1).all() and np.isinf(dists).all()
