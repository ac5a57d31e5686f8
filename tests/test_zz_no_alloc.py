"""hnswgpu_search_batch_device_noalloc (no reference counterpart): the search
of hnswgpu_search_batch_device with the library's static scratch. Through the
C ABI, two calls in a row (the second smaller) must return exactly the rows and
counts of hnswgpu_search_batch_device on the same inputs: plain, with a label
filter, with soft deletions, cosine (the device-side query scale), and a size
beyond the static scratch is refused with EINVAL."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def rows_equal(got, want):
    return np.array_equal(got["label"], want["label"]) and np.array_equal(
        got["distance"].view(np.uint32), want["distance"].view(np.uint32))


def run_both(lib, dev, Q, k, ef, filt=None):
    stream = lib.Stream()
    nq = Q.shape[0]
    d_q = lib.DeviceBuffer.of(np.ascontiguousarray(Q, np.float32))
    d_a = lib.DeviceBuffer(nq * k * lib.NEIGHBOR.itemsize)
    d_b = lib.DeviceBuffer(nq * k * lib.NEIGHBOR.itemsize)
    c_a, c_b = lib.DeviceBuffer(4 * nq), lib.DeviceBuffer(4 * nq)
    d_f, fsize = 0, 0
    if filt is not None:
        w, fsize = lib.filter_words(filt)
        fbuf = lib.DeviceBuffer.of(w)
        d_f = fbuf.ptr
    dev.search_device(d_q.ptr, nq, k, ef, d_a.ptr, stream=stream.handle, d_filter=d_f, filter_size=fsize,
                      d_counts=c_a.ptr)
    dev.search_device_noalloc(d_q.ptr, nq, k, ef, d_b.ptr, stream=stream.handle, d_filter=d_f, filter_size=fsize,
                              d_counts=c_b.ptr)
    stream.synchronize()
    a = d_a.download(np.empty(nq * k, lib.NEIGHBOR)).reshape(nq, k)
    b = d_b.download(np.empty(nq * k, lib.NEIGHBOR)).reshape(nq, k)
    ca = c_a.download(np.empty(nq, np.uint32))
    cb = c_b.download(np.empty(nq, np.uint32))
    return a, b, ca, cb


def test_noalloc_equals_device_entry_two_calls(lib):
    import workloads

    X = workloads.mixture_128(20000, seed=61)
    Q = workloads.mixture_128(3000, seed=1061)
    ix = lib.Index(128, "l2")
    ix.add(X, num_threads=0)
    dev = ix.device()
    rng = np.random.default_rng(6)
    allow = rng.random(20000) < 0.1
    for deleted in (False, True):
        if deleted:
            for lab in rng.choice(20000, 300, replace=False).tolist():
                ix.mark_deleted(lab)
        for filt in (None, allow):
            # Two calls in a row, the second with a smaller input.
            for nq, k, ef in ((3000, 10, 128), (700, 5, 64)):
                case = (deleted, filt is not None, nq, k, ef)
                print("case", case, flush=True)
                a, b, ca, cb = run_both(lib, dev, Q[:nq], k, ef, filt)
                assert rows_equal(b, a), case
                assert np.array_equal(cb, ca), case


def test_noalloc_cosine_and_large_ef(golden, lib):
    ix = lib.Index.load(GOLDEN / "cos_d100_avx512.idx")
    Q = golden["cos_d100/queries"]
    dev = ix.device()
    for nq, k, ef in ((Q.shape[0], 25, 600), (7, 10, 64)):
        a, b, ca, cb = run_both(lib, dev, Q[:nq], k, ef)
        assert rows_equal(b, a), (nq, k, ef)
        assert np.array_equal(cb, ca)
    assert rows_equal(b, golden["cos_d100/avx512/plain/k10_ef64"][:7])


def test_noalloc_refuses_sizes_beyond_its_scratch(lib):
    import workloads

    X = workloads.mixture_128(5000, seed=62)
    ix = lib.Index(128, "l2")
    ix.add(X, num_threads=0)
    dev = ix.device()
    nq = 200_000  # 200,000 x 128 prepared queries alone need 100 MiB
    d_q = lib.DeviceBuffer(nq * 128 * 4)
    d_out = lib.DeviceBuffer(nq * 10 * lib.NEIGHBOR.itemsize)
    with pytest.raises(ValueError, match="static scratch holds"):
        dev.search_device_noalloc(d_q.ptr, nq, 10, 64, d_out.ptr)
