"""Host-side checks that need no GPU.

* The cosine query scale of device-buffer searches is computed without FP64,
  by integer emulation of the reference's double arithmetic
  (src/space.hpp:31-40: a sequential double sum of squares, then
  float(1 / sqrt)). The same code runs on the host through
  hnswgpu_query_scale; here it is compared bit for bit with that arithmetic
  done in IEEE doubles (numpy / Python floats), on random vectors of every
  magnitude and on cases built to hit rounding ties, sequential-order effects,
  subnormals, overflow and NaN / infinity.
* The metric names the binding accepts are exactly those the reference's
  parse_metric accepts (src/dispatch.cpp:15-20), checked against the compiled
  reference.
* The Python module's surface (properties, membership, repr, simd_isa) passes
  the reference's own basic-properties test on a host-built index.
"""

from __future__ import annotations

import math

import numpy as np
import pytest


def scale_by_doubles(x: np.ndarray) -> np.uint32:
    """float(1 / sqrt(sum in order of double(x)^2)), 1 for a zero or NaN sum."""
    s = 0.0
    for v in x.astype(np.float64).tolist():
        s += v * v
    if not s > 0.0:
        return np.float32(1.0).view(np.uint32)
    return np.float32(1.0 / math.sqrt(s)).view(np.uint32)


def scales_by_doubles(X: np.ndarray) -> np.ndarray:
    """scale_by_doubles for every row of X, vectorised across rows (still in order along each row)."""
    s = np.zeros(X.shape[0], np.float64)
    for j in range(X.shape[1]):
        c = X[:, j].astype(np.float64)
        s = s + c * c
    with np.errstate(divide="ignore", invalid="ignore"):
        out = (1.0 / np.sqrt(s)).astype(np.float32)
    out[~(s > 0.0)] = np.float32(1.0)
    return out.view(np.uint32)


def device_scale_bits(lib, x) -> int:
    return int(np.float32(lib.query_scale(np.asarray(x, np.float32))).view(np.uint32))


def test_scale_special_cases(lib):
    f32 = np.float32
    cases = [
        [0.0] * 5,                                   # zero vector -> 1
        [1.0, 2.0 ** -27],                           # 1 + 2^-54: a tie, rounds to even (1)
        [1.0, 2.0 ** -27, 2.0 ** -27],               # in order: 1 both times (exact sum would round up)
        [2.0 ** -27, 2.0 ** -27, 1.0],               # other order: 2^-53 + 1 rounds up
        [1.0 + 2.0 ** -23, 2.0 ** -27],              # tie with an odd last bit
        [3.0, 4.0],                                  # 1/5
        [1e-45],                                     # the smallest subnormal float: 1/x overflows
        [1e-40, 3e-41],                              # subnormals
        [3.4e38],                                    # the largest floats
        [3.4e38] * 4096,                             # a scale in the subnormal range
        [-2.5, 0.0, 7.25],
        [float("nan"), 1.0],                         # NaN sum: norm > 0 is false -> 1
        [float("inf"), 1.0],                         # infinite sum -> 0
        [1.0, float("-inf")],
        [float("inf"), float("nan")],
    ]
    for c in cases:
        x = np.array(c, f32)
        assert device_scale_bits(lib, x) == int(scale_by_doubles(x)), c


@pytest.mark.parametrize("dim", [1, 3, 16, 100, 128, 960, 4096])
def test_scale_random_vectors(lib, dim):
    rng = np.random.default_rng(dim)
    n = max(50, 20000 // dim)
    mags = np.exp2(rng.integers(-140, 90, n)).astype(np.float64)
    X = (rng.standard_normal((n, dim)) * mags[:, None]).astype(np.float32)
    # Mixed magnitudes inside one vector make additions lose low bits (sticky cases).
    X[::3] *= np.exp2(rng.integers(-20, 20, (X[::3].shape[0], dim))).astype(np.float32)
    X[::5, ::2] = 0.0
    want = scales_by_doubles(X)
    got = np.array([device_scale_bits(lib, x) for x in X], np.uint32)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, (dim, bad[:5], got[bad[:5]], want[bad[:5]])


def test_scale_integer_vectors(lib):
    """Small integers (exact sums) and near-ties from integer data."""
    rng = np.random.default_rng(5)
    X = rng.integers(0, 256, (3000, 128)).astype(np.float32)
    want = scales_by_doubles(X)
    got = np.array([device_scale_bits(lib, x) for x in X], np.uint32)
    assert np.array_equal(got, want)


def test_metric_names_match_the_reference(lib, reference):
    names = ["l2", "L2", "euclidean", "ip", "inner_product", "dot", "cosine", "angular",
             "IP", "Cosine", "l1", "hamming", "", "cos", "L2 "]
    for name in names:
        ref_id = reference.parse_metric(name)
        ours = lib.METRICS.get(name, -1)
        assert ours == ref_id, (name, ours, ref_id)


def test_module_surface_matches_the_reference(lib, tmp_path):
    """The reference's test_basic_properties (tests/py/test_index.py:67-76) on
    a host-built index: sizes, properties, membership, repr and simd_isa."""
    import workloads

    X = workloads.uniform(3000, 32, seed=3)
    index = lib.Index(32, "l2", M=16, ef_construction=200)
    index.add(X, num_threads=2)
    assert len(index) == 3000
    assert index.dim == 32
    assert index.metric == "l2"
    assert index.M == 16
    assert index.ef_construction == 200
    assert not index.read_only
    assert 0 in index and 2999 in index and 3000 not in index
    assert "Index dim=32" in repr(index) and "size=3000" in repr(index)
    assert lib.simd_isa() in {"scalar", "avx2", "avx512", "neon"}
    assert isinstance(lib.__version__, str)
    path = tmp_path / "ro.idx"
    index.save(path)
    ro = lib.Index.load(path, mmap=True)
    assert ro.read_only and repr(ro).endswith(" read_only>")
