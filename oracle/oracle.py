"""TEST INFRASTRUCTURE ONLY: ctypes views of the two checkers.

* ``Oracle``     -- oracle/liboracle.so, the project's CPU restatement of the
                    reference's query path (oracle/oracle.cpp);
* ``Reference``  -- oracle/_ref/libhnswref.so, the reference itself compiled
                    from its sources plus oracle/ref_shim.cpp (portable flags:
                    the checker), or oracle/_ref/libhnswref_bench.so (the same
                    sources at -O3 -march=x86-64-v4: the CPU baseline).

Only tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py use
this module.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "liboracle.so"
REF_SO = HERE / "_ref" / "libhnswref.so"
REF_BENCH_SO = HERE / "_ref" / "libhnswref_bench.so"
ROW = np.dtype([("label", "<u8"), ("distance", "<f4"), ("reserved", "<u4")])
ORDER = {"scalar": 0, "avx2": 1, "avx512": 2}
p, u64, i64 = C.c_void_p, C.c_uint64, C.c_int64


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Oracle:
    def __init__(self):
        if not ORACLE_SO.exists():
            raise FileNotFoundError(f"{ORACLE_SO} missing: make -C oracle")
        lib = C.CDLL(str(ORACLE_SO))
        lib.orc_last_error.restype = C.c_char_p
        for f in (lib.orc_l2, lib.orc_dot):
            f.restype, f.argtypes = C.c_float, [p, p, u64, C.c_int]
        lib.orc_dist.restype, lib.orc_dist.argtypes = C.c_float, [C.c_int, p, p, u64, C.c_int]
        lib.orc_prepare.argtypes = [C.c_int, p, p, u64]
        lib.orc_load.argtypes = [C.c_char_p, p]
        lib.orc_free.argtypes = [p]
        lib.orc_info.argtypes = [p, p]
        lib.orc_mark_deleted.argtypes = [p, u64]
        lib.orc_search_batch.argtypes = [p, p, u64, u64, u64, C.c_int, p, u64, p, p, u64]
        lib.orc_brute_force.argtypes = [u64, C.c_int, p, p, u64, p, u64, u64, C.c_int, p, u64, p, u64]
        lib.orc_label_dists.argtypes = [p, p, u64, p, u64, C.c_int, p, u64]
        self.lib = lib

    def _chk(self, rc):
        if rc != 0:
            raise RuntimeError(self.lib.orc_last_error().decode())

    def l2(self, a, b, order="avx512"):
        a = np.ascontiguousarray(a, np.float32); b = np.ascontiguousarray(b, np.float32)
        return self.lib.orc_l2(_ptr(a), _ptr(b), a.size, ORDER[order])

    def dot(self, a, b, order="avx512"):
        a = np.ascontiguousarray(a, np.float32); b = np.ascontiguousarray(b, np.float32)
        return self.lib.orc_dot(_ptr(a), _ptr(b), a.size, ORDER[order])

    def load(self, path):
        h = C.c_void_p()
        self._chk(self.lib.orc_load(os.fsencode(path), C.byref(h)))
        return OracleIndex(self, h)

    def brute_force(self, data, queries, k, metric=0, labels=None, order="avx512", filter=None, threads=0):
        d = np.ascontiguousarray(data, np.float32); q = np.ascontiguousarray(queries, np.float32)
        lab = None if labels is None else np.ascontiguousarray(labels, np.uint64)
        w, size = (None, 0) if filter is None else filter
        out = np.empty(q.shape[0] * k, ROW)
        self._chk(self.lib.orc_brute_force(d.shape[1], metric, _ptr(d), _ptr(lab), d.shape[0], _ptr(q), q.shape[0],
                                           k, ORDER[order], _ptr(w), size, _ptr(out), threads))
        return out.reshape(q.shape[0], k)


class OracleIndex:
    def __init__(self, orc, h):
        self.orc, self.h = orc, h

    def __del__(self):
        if getattr(self, "h", None):
            self.orc.lib.orc_free(self.h)
            self.h = None

    def info(self):
        a = np.zeros(9, np.int64)
        self.orc.lib.orc_info(self.h, _ptr(a))
        keys = ["dim", "padded", "M", "count", "max_level", "entry", "metric", "ef_search", "num_deleted"]
        return dict(zip(keys, a.tolist()))

    def mark_deleted(self, label):
        self.orc._chk(self.orc.lib.orc_mark_deleted(self.h, label))

    def search(self, queries, k, ef, order="avx512", filter=None, counters=False, threads=0):
        q = np.ascontiguousarray(queries, np.float32)
        nq = q.shape[0]
        out = np.empty(nq * k, ROW)
        cnt = np.zeros(nq * 3, np.uint64) if counters else None
        w, size = (None, 0) if filter is None else filter
        self.orc._chk(self.orc.lib.orc_search_batch(self.h, _ptr(q), nq, k, ef, ORDER[order], _ptr(w), size,
                                                    _ptr(out), _ptr(cnt), threads))
        out = out.reshape(nq, k)
        return (out, cnt.reshape(nq, 3)) if counters else out


def _label_dists(self, queries, labels, order="avx512", threads=0):
    """Exact distance (this summation order) of query i to the node of each
    label in labels[i]; +inf for invalid labels."""
    q = np.ascontiguousarray(queries, np.float32)
    lab = np.ascontiguousarray(labels, np.uint64)
    out = np.empty(lab.shape, np.float32)
    self.orc._chk(self.orc.lib.orc_label_dists(self.h, _ptr(q), q.shape[0], _ptr(lab), lab.shape[1], ORDER[order],
                                               _ptr(out), threads))
    return out


OracleIndex.label_dists = _label_dists


class Reference:
    """The reference library compiled from its own sources (oracle/_ref)."""

    def __init__(self, so=REF_SO):
        if not Path(so).exists():
            raise FileNotFoundError(f"{so} missing: make -C oracle ref (needs the reference sources)")
        lib = C.CDLL(str(so))
        lib.ref_last_error.restype = C.c_char_p
        lib.ref_set_isa.argtypes = [C.c_char_p]
        lib.ref_active_isa.argtypes = [p, C.c_int]
        lib.ref_build_save.argtypes = [C.c_char_p, u64, C.c_int, u64, u64, u64, p, p, u64, u64]
        lib.ref_load.argtypes = [C.c_char_p, C.c_int, p]
        lib.ref_free.argtypes = [p]
        lib.ref_mark_deleted.argtypes = [p, u64]
        lib.ref_set_ef.argtypes = [p, u64]
        lib.ref_save.argtypes = [p, C.c_char_p]
        lib.ref_search_batch.argtypes = [p, p, u64, u64, u64, p, u64, p, u64]
        lib.ref_brute_force.argtypes = [u64, C.c_int, p, p, u64, p, u64, u64, p, u64, p, u64]
        lib.ref_parse_metric.argtypes = [C.c_char_p]
        for f in (lib.ref_l2_sq, lib.ref_dot):
            f.restype, f.argtypes = C.c_float, [p, p, u64]
        self.lib = lib

    def _chk(self, rc):
        if rc != 0:
            raise RuntimeError(self.lib.ref_last_error().decode())

    def set_isa(self, name):
        self._chk(self.lib.ref_set_isa(name.encode()))

    def parse_metric(self, name):
        """The reference's parse_metric: the metric id, or -1 if it throws."""
        return self.lib.ref_parse_metric(name.encode())

    def active_isa(self):
        b = C.create_string_buffer(32)
        self.lib.ref_active_isa(b, 32)
        return b.value.decode()

    def build_save(self, path, data, labels=None, metric=0, M=16, efc=200, seed=100, threads=1):
        d = np.ascontiguousarray(data, np.float32)
        lab = np.arange(d.shape[0], dtype=np.uint64) if labels is None else np.ascontiguousarray(labels, np.uint64)
        self._chk(self.lib.ref_build_save(os.fsencode(path), d.shape[1], metric, M, efc, seed, _ptr(d), _ptr(lab),
                                          d.shape[0], threads))

    def load(self, path, mmap=False):
        h = C.c_void_p()
        self._chk(self.lib.ref_load(os.fsencode(path), 1 if mmap else 0, C.byref(h)))
        return RefIndex(self, h)

    def brute_force(self, data, queries, k, metric=0, labels=None, filter=None, threads=0):
        d = np.ascontiguousarray(data, np.float32); q = np.ascontiguousarray(queries, np.float32)
        lab = None if labels is None else np.ascontiguousarray(labels, np.uint64)
        w, size = (None, 0) if filter is None else filter
        out = np.empty(q.shape[0] * k, ROW)
        self._chk(self.lib.ref_brute_force(d.shape[1], metric, _ptr(d), _ptr(lab), d.shape[0], _ptr(q), q.shape[0],
                                           k, _ptr(out), threads, _ptr(w), size))
        return out.reshape(q.shape[0], k)


class RefIndex:
    def __init__(self, ref, h):
        self.ref, self.h = ref, h

    def __del__(self):
        if getattr(self, "h", None):
            self.ref.lib.ref_free(self.h)
            self.h = None

    def mark_deleted(self, label):
        self.ref._chk(self.ref.lib.ref_mark_deleted(self.h, label))

    def set_ef(self, ef):
        self.ref._chk(self.ref.lib.ref_set_ef(self.h, ef))

    def save(self, path):
        self.ref._chk(self.ref.lib.ref_save(self.h, os.fsencode(path)))

    def search(self, queries, k, ef, threads=1, filter=None):
        q = np.ascontiguousarray(queries, np.float32)
        out = np.empty(q.shape[0] * k, ROW)
        w, size = (None, 0) if filter is None else filter
        self.ref._chk(self.ref.lib.ref_search_batch(self.h, _ptr(q), q.shape[0], k, ef, _ptr(out), threads, _ptr(w),
                                                    size))
        return out.reshape(q.shape[0], k)
