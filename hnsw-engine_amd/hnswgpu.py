"""Python view of the engine's C ABI (include/hnswgpu.h) over ctypes.

The classes mirror the reference's Python binding (python/bindings.cpp:121-233
of the reference): ``Index(dim, metric, M, ef_construction, max_elements,
seed)``, ``add``, ``search`` returning ``(ids, distances)`` with -1 / +inf
padding, ``mark_deleted``, ``save`` / ``load``; searches run on the GPU.
``DeviceIndex`` exposes the device layer (``hnswgpu_*``) for device buffers
and tuning. Nothing here computes: every call crosses into libhnswgpu.so, and
importing fails loudly when the library has not been built.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

# The serving loop keeps several launches in flight, one stream per slot. HIP
# maps streams onto a pool of hardware queues (4 per process by default) and
# launches of streams that share a queue run one after another, so with the
# index's own stream and the default stream beside three slots, two slots can
# end up serialised. Unless the caller chose a value, ask for 16 queues (room
# for 8 slots beside the search contexts' streams); the HIP runtime reads this
# once, when it starts (the first HIP call).
os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libhnswgpu.so"

METRICS = {"l2": 0, "L2": 0, "euclidean": 0, "ip": 1, "inner_product": 1, "dot": 1, "cosine": 2, "angular": 2}
METRIC_NAMES = {0: "l2", 1: "ip", 2: "cosine"}

NEIGHBOR = np.dtype([("label", "<u8"), ("distance", "<f4"), ("reserved", "<u4")])
INVALID_LABEL = np.uint64(0xFFFFFFFFFFFFFFFF)

OK, EINVAL, ERUNTIME, EDEVICE, ERANGE = 0, -1, -2, -3, -4


def _load():
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() (make -C hnsw-engine_amd)")
    lib = C.CDLL(str(LIB_PATH))
    p, u64, i32, i64 = C.c_void_p, C.c_uint64, C.c_int32, C.c_int64
    sig = {
        "hnswgpu_last_error": ([], C.c_char_p),
        "hnswgpu_upload": ([p, p], C.c_int),
        "hnswgpu_free": ([p], C.c_int),
        "hnswgpu_set_deleted": ([p, C.c_uint32, C.c_uint8], C.c_int),
        "hnswgpu_set_ef": ([p, u64], C.c_int),
        "hnswgpu_search_batch": ([p, p, u64, u64, u64, p, u64, p, p, u64], C.c_int),
        "hnswgpu_search_batch_device": ([p, p, u64, u64, u64, p, u64, p, p, p], C.c_int),
        "hnswgpu_search_batch_ids": ([p, p, u64, u64, u64, p, p, p, u64], C.c_int),
        "hnswgpu_search_batch_device_noalloc": ([p, p, u64, u64, u64, p, u64, p, p, p], C.c_int),
        "hnswgpu_search_batch_ids_device": ([p, p, u64, u64, u64, p, p, p, p], C.c_int),
        "hnswgpu_reupload": ([p, p], C.c_int),
        "hnswgpu_query_scale": ([p, u64, p], C.c_int),
        "hnsw_index_search_batch_pred": ([p, p, u64, u64, u64, p, u64, p, p], C.c_int),
        "hnswgpu_set_option": ([p, C.c_char_p, i64], C.c_int),
        "hnswgpu_get_option": ([p, C.c_char_p, p], C.c_int),
        "hnswgpu_read_stats": ([p, p, u64], C.c_int),
        "hnswgpu_get_row": ([p, C.c_uint32, C.c_int, p, p], C.c_int),
        "hnswgpu_malloc": ([u64, p], C.c_int),
        "hnswgpu_mfree": ([p], C.c_int),
        "hnswgpu_memcpy": ([p, p, u64, C.c_int], C.c_int),
        "hnswgpu_copy_kernel": ([p, p, u64], C.c_int),
        "hnswgpu_stream_create": ([p], C.c_int),
        "hnswgpu_stream_destroy": ([p], C.c_int),
        "hnswgpu_stream_synchronize": ([p], C.c_int),
        "hnswgpu_event_create": ([p], C.c_int),
        "hnswgpu_event_destroy": ([p], C.c_int),
        "hnswgpu_event_record": ([p, p], C.c_int),
        "hnswgpu_event_elapsed_ms": ([p, p, p], C.c_int),
        "hnswgpu_device_info": ([p, C.c_int, p], C.c_int),
        "hnsw_index_create": ([u64, i32, u64, u64, u64, u64, p], C.c_int),
        "hnsw_index_load": ([C.c_char_p, C.c_int, p], C.c_int),
        "hnsw_index_free": ([p], None),
        "hnsw_index_save": ([p, C.c_char_p], C.c_int),
        "hnsw_index_add_batch": ([p, p, p, u64, u64], C.c_int),
        "hnsw_index_add_batch_device": ([p, p, p, u64, u64, p], C.c_int),
        "hnsw_index_search_batch": ([p, p, u64, u64, u64, p, u64, p, u64], C.c_int),
        "hnsw_index_search": ([p, p, u64, u64, p, u64, p, p], C.c_int),
        "hnsw_index_mark_deleted": ([p, u64], C.c_int),
        "hnsw_index_is_deleted": ([p, u64, p], C.c_int),
        "hnsw_index_contains": ([p, u64, p], C.c_int),
        "hnsw_index_set_ef": ([p, u64], C.c_int),
        "hnsw_index_get_vector": ([p, u64, p], C.c_int),
        "hnsw_index_info": ([p, p], C.c_int),
        "hnsw_index_device": ([p, p], C.c_int),
        "hnsw_brute_force_create": ([u64, i32, p], C.c_int),
        "hnsw_brute_force_free": ([p], None),
        "hnsw_brute_force_add": ([p, p, p, u64], C.c_int),
        "hnsw_brute_force_search_batch": ([p, p, u64, u64, p, u64, p, u64], C.c_int),
        "hnswgpu_search_tasks_device": ([p, p, u64, p, u64, i32, p], C.c_int),
        "hnswgpu_autotune": ([p, p, u64, u64, p, u64, p, u64], C.c_int),
        "hnswgpu_server_create": ([p, p, p], C.c_int),
        "hnswgpu_server_submit": ([p, p, p], C.c_int),
        "hnswgpu_server_wait": ([p, u64], C.c_int),
        "hnswgpu_server_wait_latency": ([p, u64, p], C.c_int),
        "hnswgpu_server_stats": ([p, p, u64], C.c_int),
        "hnswgpu_server_reset_stats": ([p], C.c_int),
        "hnswgpu_server_free": ([p], C.c_int),
        "hnsw_brute_force_set_option": ([p, C.c_char_p, C.c_int64], C.c_int),
        "hnsw_brute_force_get_option": ([p, C.c_char_p, p], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    return lib


_lib = _load()


def _check(rc: int) -> None:
    if rc == OK:
        return
    msg = (_lib.hnswgpu_last_error() or b"").decode()
    if rc == EINVAL:
        raise ValueError(msg)
    if rc == ERANGE:
        raise IndexError(msg)
    raise RuntimeError(msg)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f32(x, dim, what):
    """(n, dim) C-contiguous float32 copy or view of x, and whether x was one
    vector. The reference binding's rules (python/bindings.cpp:25-48): any
    numeric input is cast, anything else is a TypeError, a wrong shape is a
    ValueError."""
    try:
        a = np.asarray(x)
    except (TypeError, ValueError):
        a = None
    if a is None or a.dtype.kind not in "biuf":
        raise TypeError(f"{what}: expected a numeric array convertible to float32")
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim == 1:
        if a.shape[0] != dim:
            raise ValueError(f"{what}: expected {dim} values, got {a.shape[0]}")
        return a.reshape(1, -1), True
    if a.ndim != 2:
        raise ValueError(f"{what}: expected a 1-D or 2-D array")
    if a.shape[1] != dim:
        raise ValueError(f"{what}: expected shape (n, {dim}), got ({a.shape[0]}, {a.shape[1]})")
    return a, False


def _labels(ids, n: int, start: int) -> np.ndarray:
    """uint64 labels of n added vectors: start, start + 1, ... for None, else
    a 1-D integer array of n non-negative labels (python/bindings.cpp:50-77 of
    the reference: wrong length ValueError, negative ValueError, non-integer
    TypeError)."""
    if ids is None:
        return np.arange(start, start + n, dtype=np.uint64)
    try:
        a = np.asarray(ids)
    except (TypeError, ValueError):
        raise TypeError("ids: expected an integer array") from None
    if a.ndim == 0:
        a = a.reshape(1)
    if a.ndim != 1 or a.shape[0] != n:
        raise ValueError(f"ids: expected shape ({n},)")
    if a.dtype.kind == "i":
        if (a < 0).any():
            raise ValueError("ids: labels must be non-negative")
    elif a.dtype.kind != "u":
        raise TypeError("ids: expected an integer array")
    return np.ascontiguousarray(a, dtype=np.uint64)


def _search_filter(filter):
    """The filter argument of Index.search as the reference binding reads it
    (python/bindings.cpp:79-108): None; a 1-D array, list or tuple that is a
    bool mask over labels or a list of allowed labels (a BitsetFilter, here
    its (words, size)); or a callable(label) -> bool (a FunctionFilter)."""
    if filter is None:
        return None
    if isinstance(filter, (np.ndarray, list, tuple)):
        try:
            a = np.asarray(filter)
        except (TypeError, ValueError):
            a = None
        if a is None or a.ndim != 1:
            raise ValueError("filter: expected a 1-D array")
        if a.dtype.kind == "b":
            return filter_words(a)
        if a.dtype.kind in "iu":
            return filter_words(_labels(a, a.shape[0], 0))
        raise TypeError("filter: array must be bool (mask over labels) or integer (allowed labels)")
    if callable(filter):
        return filter
    raise TypeError("filter: expected None, a callable(label) -> bool, a bool mask or an id array")


LABEL_PREDICATE = C.CFUNCTYPE(C.c_int, C.c_uint64, C.c_void_p)


def query_scale(x) -> float:
    """The cosine query scale 1/|x| the device computes for device-buffer
    searches (integer emulation of the reference's double arithmetic), run on
    the host (hnswgpu_query_scale)."""
    a = np.ascontiguousarray(x, dtype=np.float32).reshape(-1)
    out = np.zeros(1, np.float32)
    _check(_lib.hnswgpu_query_scale(_ptr(a), a.size, _ptr(out)))
    return out[0]


def id_words(mask) -> np.ndarray:
    """Words of a bitmap over internal ids from a bool mask (hnswgpu_search_batch_ids)."""
    m = np.asarray(mask, dtype=bool).reshape(-1)
    bits = np.zeros(max(1, (m.size + 63) // 64) * 64, dtype=np.uint64)
    bits[:m.size] = m
    return np.ascontiguousarray(np.bitwise_or.reduce(bits.reshape(-1, 64) << np.arange(64, dtype=np.uint64), axis=1))


def filter_words(filter, num_labels=None):
    """(words, size) of a label bitset from a bool mask over labels, an array of
    allowed labels, or an existing (words, size) pair (None -> (None, 0)).
    Callables are not label bitsets: Index.search takes them as predicates."""
    if filter is None:
        return None, 0
    if callable(filter):
        raise TypeError("filter: a callable is a predicate, not a label bitset (use Index.search / search_rows)")
    if isinstance(filter, tuple):
        w, size = filter
        return np.ascontiguousarray(w, dtype=np.uint64), int(size)
    a = np.asarray(filter)
    if a.dtype == np.bool_:
        size = a.size
        bits = np.zeros(max(1, (size + 63) // 64) * 64, dtype=np.uint64)
        bits[:size] = a.reshape(-1)
        words = np.bitwise_or.reduce(bits.reshape(-1, 64) << np.arange(64, dtype=np.uint64), axis=1)
        return np.ascontiguousarray(words, dtype=np.uint64), size
    labels = a.astype(np.uint64).reshape(-1)
    size = int(labels.max()) + 1 if labels.size else 0
    words = np.zeros(max(1, (size + 63) // 64), dtype=np.uint64)
    np.bitwise_or.at(words, (labels >> np.uint64(6)).astype(np.int64), np.uint64(1) << (labels & np.uint64(63)))
    return words, size


class DeviceBuffer:
    """A device allocation of the library's runtime (hnswgpu_malloc)."""

    def __init__(self, nbytes: int):
        h = C.c_void_p()
        _check(_lib.hnswgpu_malloc(max(1, int(nbytes)), C.byref(h)))
        self.ptr, self.nbytes = h.value, int(nbytes)

    @staticmethod
    def of(array: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(array)
        b = DeviceBuffer(a.nbytes)
        b.upload(a)
        return b

    def upload(self, array: np.ndarray) -> None:
        a = np.ascontiguousarray(array)
        assert a.nbytes <= self.nbytes
        _check(_lib.hnswgpu_memcpy(self.ptr, _ptr(a), a.nbytes, 0))

    def download(self, out: np.ndarray) -> np.ndarray:
        assert out.flags.c_contiguous and out.nbytes <= self.nbytes
        _check(_lib.hnswgpu_memcpy(_ptr(out), self.ptr, out.nbytes, 1))
        return out

    def __del__(self):
        if getattr(self, "ptr", None):
            _lib.hnswgpu_mfree(self.ptr)
            self.ptr = None


class Stream:
    def __init__(self):
        h = C.c_void_p()
        _check(_lib.hnswgpu_stream_create(C.byref(h)))
        self.handle = h.value

    def synchronize(self) -> None:
        _check(_lib.hnswgpu_stream_synchronize(self.handle))

    def __del__(self):
        if getattr(self, "handle", None):
            _lib.hnswgpu_stream_destroy(self.handle)
            self.handle = None


class Event:
    def __init__(self):
        h = C.c_void_p()
        _check(_lib.hnswgpu_event_create(C.byref(h)))
        self.handle = h.value

    def record(self, stream: "Stream") -> None:
        _check(_lib.hnswgpu_event_record(self.handle, stream.handle))

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        _check(_lib.hnswgpu_event_elapsed_ms(self.handle, end.handle, C.byref(ms)))
        return ms.value

    def __del__(self):
        if getattr(self, "handle", None):
            _lib.hnswgpu_event_destroy(self.handle)
            self.handle = None


def copy_device(dst: DeviceBuffer, src: DeviceBuffer, nbytes: int) -> None:
    """Copy with the library's plain copy kernel (synchronous)."""
    _check(_lib.hnswgpu_copy_kernel(dst.ptr, src.ptr, int(nbytes)))


def device_info():
    name = C.create_string_buffer(128)
    info = np.zeros(6, np.int64)
    _check(_lib.hnswgpu_device_info(name, 128, _ptr(info)))
    keys = ["cus", "global_mem", "lds_per_cu", "clock_khz", "mem_clock_khz", "mem_bus_width"]
    return {"arch": name.value.decode(), **dict(zip(keys, info.tolist()))}


class Task(C.Structure):
    """hnswgpu_task: one query of a mixed launch (device pointers)."""
    _fields_ = [("query", C.c_void_p), ("out", C.c_void_p), ("count", C.c_void_p), ("k", C.c_uint32),
                ("ef", C.c_uint32), ("filter", C.c_int32), ("reserved", C.c_int32)]


class FilterRef(C.Structure):
    """hnswgpu_filter: BitsetFilter words on the device and their label count."""
    _fields_ = [("words", C.c_void_p), ("size", C.c_uint64)]


class ServerConfig(C.Structure):
    _fields_ = [("slots", C.c_uint32), ("max_batch", C.c_uint32), ("max_rows", C.c_uint64),
                ("max_filters", C.c_uint32), ("use_graphs", C.c_uint32), ("max_filter_labels", C.c_uint64),
                ("hbm_budget", C.c_uint64), ("spill_waves", C.c_uint32), ("storages", C.c_uint32),
                ("max_cost", C.c_uint64), ("heavy_cost", C.c_uint64)]


class BuildStats(C.Structure):
    """hnswgpu_build_stats: what a device-probed add did."""
    _fields_ = [("waves", C.c_uint64), ("device_nodes", C.c_uint64), ("host_nodes", C.c_uint64),
                ("device_ms", C.c_double), ("refresh_ms", C.c_double)]


class Request(C.Structure):
    _fields_ = [("queries", C.c_void_p), ("nq", C.c_uint64), ("k", C.c_uint64), ("ef", C.c_uint64),
                ("filter_words", C.c_void_p), ("filter_size", C.c_uint64), ("storage", C.c_int32),
                ("reserved", C.c_int32), ("out", C.c_void_p), ("counts", C.c_void_p)]


SERVER_STATS = ["requests", "queries", "launches", "mean_batch", "p50_ms", "p99_ms", "max_ms", "qps", "mean_ms",
                "slots", "slot_bytes", "graph_launches", "recaptures"]


class Server:
    """Continuous-batching serving loop over a resident index (hnswgpu_server_*).
    submit() queues a request and returns a ticket; wait(ticket) returns its
    (rows, counts), equal to search_batch on the same arguments."""

    def __init__(self, device: "DeviceIndex", slots: int = 2, max_batch: int = 4096, max_rows: int = 0,
                 max_filters: int = 8, graphs: bool = True, max_filter_labels: int = 0, hbm_budget: int = 0,
                 spill_waves: int = 0, storages=(), max_cost: int = 0, heavy_cost: int = 0):
        cfg = ServerConfig(slots, max_batch, max_rows, max_filters, 1 if graphs else 0, max_filter_labels,
                           hbm_budget, spill_waves, sum(1 << int(st) for st in storages), max_cost, heavy_cost)
        h = C.c_void_p()
        _check(_lib.hnswgpu_server_create(device._h, C.byref(cfg), C.byref(h)))
        self._h, self._device, self._live = h, device, {}

    def submit(self, queries, k: int, ef: int = 0, filter=None, storage: int = -1) -> int:
        q = np.ascontiguousarray(queries, dtype=np.float32)
        if q.ndim == 1:
            q = q.reshape(1, -1)
        nq = q.shape[0]
        out = np.empty(nq * k, dtype=NEIGHBOR)
        counts = np.empty(nq, dtype=np.uint32)
        w, size = filter_words(filter)
        req = Request(_ptr(q), nq, k, ef, _ptr(w), size, storage, 0, _ptr(out), _ptr(counts))
        t = C.c_uint64()
        _check(_lib.hnswgpu_server_submit(self._h, C.byref(req), C.byref(t)))
        self._live[t.value] = (q, w, out, counts)  # buffers stay alive until wait()
        return t.value

    def wait(self, ticket: int):
        rc = _lib.hnswgpu_server_wait(self._h, ticket)
        q, w, out, counts = self._live.pop(ticket)
        _check(rc)
        return out.reshape(q.shape[0], -1), counts

    def wait_latency(self, ticket: int) -> float:
        """wait() returning the request's latency (ms, submit to completion)
        instead of its rows (which are in the buffers submit made)."""
        ms = C.c_double()
        rc = _lib.hnswgpu_server_wait_latency(self._h, ticket, C.byref(ms))
        self._live.pop(ticket)
        _check(rc)
        return ms.value

    def stats(self) -> dict:
        a = np.zeros(len(SERVER_STATS), np.float64)
        _check(_lib.hnswgpu_server_stats(self._h, _ptr(a), a.size))
        return dict(zip(SERVER_STATS, a.tolist()))

    def reset_stats(self) -> None:
        _check(_lib.hnswgpu_server_reset_stats(self._h))

    def close(self) -> None:
        h = getattr(self, "_h", None)
        if h:
            self._h = None
            _check(_lib.hnswgpu_server_free(h))
            self._live.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceIndex:
    """The device layer of an index (hnswgpu_*). Owned by its Index."""

    def __init__(self, handle, owner=None):
        self._h = handle
        self._owner = owner

    def set_option(self, name: str, value: int) -> None:
        _check(_lib.hnswgpu_set_option(self._h, name.encode(), int(value)))

    def get_option(self, name: str) -> int:
        v = C.c_int64()
        _check(_lib.hnswgpu_get_option(self._h, name.encode(), C.byref(v)))
        return v.value

    def search_rows(self, queries, k, ef=0, filter=None, dim=None):
        q = np.ascontiguousarray(queries, dtype=np.float32)
        nq = q.shape[0]
        out = np.empty(nq * k, dtype=NEIGHBOR)
        counts = np.empty(nq, dtype=np.uint32)
        w, size = filter_words(filter)
        _check(_lib.hnswgpu_search_batch(self._h, _ptr(q), nq, k, ef, _ptr(w), size, _ptr(out), _ptr(counts), 0))
        return out.reshape(nq, k), counts

    def search_rows_ids(self, queries, k, ef=0, id_mask=None):
        """search_batch with an eligibility bitmap over internal ids
        (hnswgpu_search_batch_ids): the GPU form of any Filter."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        nq = q.shape[0]
        out = np.empty(nq * k, dtype=NEIGHBOR)
        counts = np.empty(nq, dtype=np.uint32)
        w = None if id_mask is None else id_words(id_mask)
        _check(_lib.hnswgpu_search_batch_ids(self._h, _ptr(q), nq, k, ef, _ptr(w), _ptr(out), _ptr(counts), 0))
        return out.reshape(nq, k), counts

    def get_row(self, node_id: int, storage: int, padded_dim: int):
        """The stored row of an internal id: (values, scale, |x^|^2)."""
        dtype = {0: np.float32, 1: np.uint16, 2: np.uint8}[storage]
        out = np.empty(padded_dim, dtype=dtype)
        sn = np.zeros(2, np.float32)
        _check(_lib.hnswgpu_get_row(self._h, node_id, storage, _ptr(out), _ptr(sn)))
        return out, float(sn[0]), float(sn[1])

    def read_stats(self, nq: int):
        out = np.empty((nq, 3), dtype=np.uint32)
        _check(_lib.hnswgpu_read_stats(self._h, _ptr(out), nq))
        return out

    def autotune(self, queries, k: int, efs) -> str:
        """Time every launch shape per ef class on these queries and keep the
        fastest (hnswgpu_autotune); returns the measurement report."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        e = np.ascontiguousarray(efs, dtype=np.uint64)
        buf = C.create_string_buffer(1 << 16)
        _check(_lib.hnswgpu_autotune(self._h, _ptr(q), q.shape[0], k, _ptr(e), e.size, buf, len(buf)))
        return buf.value.decode()

    def search_mixed(self, queries, ks, efs, filter_ids=None, filters=(), storage: int = -1):
        """One launch for queries with their own k, ef and filter (a filter id
        indexes `filters`, -1 = none). Returns a list of (rows, count)."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        n, dim = q.shape
        ks = [int(k) for k in ks]
        efs = [int(e) for e in efs]
        fids = [-1] * n if filter_ids is None else [int(f) for f in filter_ids]
        d_q = DeviceBuffer.of(q)
        offs = np.concatenate([[0], np.cumsum(ks)]).astype(np.int64)
        d_out = DeviceBuffer(int(offs[-1]) * NEIGHBOR.itemsize)
        d_cnt = DeviceBuffer(4 * n)
        fbufs, frefs = [], (FilterRef * max(1, len(filters)))()
        for i, f in enumerate(filters):
            w, size = filter_words(f)
            b = DeviceBuffer.of(w if w.size else np.zeros(1, np.uint64))
            fbufs.append(b)
            frefs[i] = FilterRef(b.ptr, size)
        tasks = (Task * n)()
        for i in range(n):
            tasks[i] = Task(d_q.ptr + i * dim * 4, d_out.ptr + int(offs[i]) * NEIGHBOR.itemsize, d_cnt.ptr + 4 * i,
                            ks[i], efs[i], fids[i], 0)
        stream = Stream()
        _check(_lib.hnswgpu_search_tasks_device(self._h, C.cast(tasks, C.c_void_p), n, C.cast(frefs, C.c_void_p),
                                                len(filters), storage, stream.handle))
        stream.synchronize()
        rows = d_out.download(np.empty(int(offs[-1]), NEIGHBOR))
        counts = d_cnt.download(np.empty(n, np.uint32))
        return [(rows[offs[i]:offs[i + 1]], int(counts[i])) for i in range(n)]

    def search_device(self, d_queries: int, nq: int, k: int, ef: int, d_out: int, stream: int = 0,
                      d_filter: int = 0, filter_size: int = 0, d_counts: int = 0) -> None:
        _check(_lib.hnswgpu_search_batch_device(self._h, d_queries, nq, k, ef, d_filter or None, filter_size,
                                                d_out, d_counts or None, stream or None))


    def search_device_noalloc(self, d_queries: int, nq: int, k: int, ef: int, d_out: int, stream: int = 0,
                              d_filter: int = 0, filter_size: int = 0, d_counts: int = 0) -> None:
        """search_device with the library's static scratch (allocates nothing)."""
        _check(_lib.hnswgpu_search_batch_device_noalloc(self._h, d_queries, nq, k, ef, d_filter or None, filter_size,
                                                        d_out, d_counts or None, stream or None))


__version__ = "0.1.0"


def simd_isa() -> str:
    """Name of the reference kernel whose fp32 summation order every distance
    of this library follows (the reference's simd_isa, python/bindings.cpp:117-119):
    always "avx512", in the GPU kernels and in the host builder alike, whatever
    the host CPU supports."""
    return "avx512"


class Index:
    """hnsw::Index over the C ABI."""

    def __init__(self, dim: int, metric: str = "l2", M: int = 16, ef_construction: int = 200,
                 max_elements: int = 0, seed: int = 100, *, _handle=None):
        if _handle is not None:
            self._h = _handle
            return
        if metric not in METRICS:
            raise ValueError(f"unknown metric '{metric}' (expected l2, ip or cosine)")
        h = C.c_void_p()
        _check(_lib.hnsw_index_create(dim, METRICS[metric], M, ef_construction, max_elements, seed, C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.hnsw_index_free(h)
            self._h = None

    @staticmethod
    def load(path, mmap: bool = False) -> "Index":
        h = C.c_void_p()
        _check(_lib.hnsw_index_load(os.fsencode(path), 1 if mmap else 0, C.byref(h)))
        return Index(0, _handle=h)

    def save(self, path) -> None:
        _check(_lib.hnsw_index_save(self._h, os.fsencode(path)))

    def _info(self):
        a = np.zeros(10, dtype=np.int64)
        _check(_lib.hnsw_index_info(self._h, _ptr(a)))
        return a

    def add(self, vectors, ids=None, num_threads: int = 0, device: bool = False):
        """add_batch; device=True runs the insertion searches of level-0 nodes
        on the GPU (the same graph as the host build with the same
        num_threads) and returns its statistics as a dict."""
        v, _ = _f32(vectors, self.dim, "vectors")
        n = v.shape[0]
        labels = _labels(ids, n, len(self))
        if not device:
            _check(_lib.hnsw_index_add_batch(self._h, _ptr(v), _ptr(labels), n, num_threads))
            return None
        st = BuildStats()
        _check(_lib.hnsw_index_add_batch_device(self._h, _ptr(v), _ptr(labels), n, num_threads, C.byref(st)))
        return {f: getattr(st, f) for f, _ in BuildStats._fields_}

    def search_rows(self, queries, k: int = 10, ef=None, filter=None):
        """Raw Neighbor rows (nq, k) with fields label/distance. filter: None,
        a bool mask over labels or an array of allowed labels (a BitsetFilter),
        or a callable(label) -> bool (a FunctionFilter, called once per node,
        as python/bindings.cpp:101-107 of the reference wraps callables)."""
        if k <= 0:
            raise ValueError("k must be > 0")
        q, _ = _f32(queries, self.dim, "queries")
        out = np.empty(q.shape[0] * k, dtype=NEIGHBOR)
        if callable(filter):
            raised = []

            def allowed(label, _ctx):
                if raised:
                    return 0
                try:
                    return 1 if filter(int(label)) else 0
                except BaseException as e:  # re-raised below, as the binding propagates it
                    raised.append(e)
                    return 0

            cb = LABEL_PREDICATE(allowed)
            rc = _lib.hnsw_index_search_batch_pred(self._h, _ptr(q), q.shape[0], k, 0 if ef is None else ef,
                                                   _ptr(out), 0, cb, None)
            if raised:
                raise raised[0]
            _check(rc)
            return out.reshape(q.shape[0], k)
        w, size = filter_words(filter)
        _check(_lib.hnsw_index_search_batch(self._h, _ptr(q), q.shape[0], k, 0 if ef is None else ef, _ptr(out),
                                            0, _ptr(w), size))
        return out.reshape(q.shape[0], k)

    def search(self, queries, k: int = 10, ef=None, num_threads: int = 0, filter=None):
        """(ids int64, distances float32), (nq, k) or (k,) for one vector; -1 /
        inf where fewer than k results exist. num_threads is accepted and not
        used: the GPU runs the whole batch."""
        if k <= 0:
            raise ValueError("k must be > 0")
        q, single = _f32(queries, self.dim, "queries")
        rows = self.search_rows(q, k, ef, _search_filter(filter))
        ids = rows["label"].astype(np.int64)  # kInvalidLabel -> -1
        d = rows["distance"].copy()
        return (ids[0], d[0]) if single else (ids, d)

    def search_one(self, query, k: int, ef: int = 0, filter=None):
        """Index::search: up to k rows (no padding)."""
        q, _ = _f32(query, self.dim, "query")
        out = np.empty(k, dtype=NEIGHBOR)
        n = C.c_uint64()
        w, size = filter_words(filter)
        _check(_lib.hnsw_index_search(self._h, _ptr(q), k, ef, _ptr(w), size, _ptr(out), C.byref(n)))
        return out[: n.value]

    def set_ef(self, ef: int) -> None:
        _check(_lib.hnsw_index_set_ef(self._h, ef))

    def mark_deleted(self, label: int) -> None:
        if label < 0:
            raise ValueError("labels are non-negative")
        _check(_lib.hnsw_index_mark_deleted(self._h, label))

    def is_deleted(self, label: int) -> bool:
        o = C.c_int()
        _check(_lib.hnsw_index_is_deleted(self._h, label, C.byref(o)))
        return bool(o.value)

    def __contains__(self, label: int) -> bool:
        o = C.c_int()
        _check(_lib.hnsw_index_contains(self._h, label, C.byref(o)))
        return bool(o.value)

    def get_vector(self, label: int):
        out = np.empty(self.dim, dtype=np.float32)
        _check(_lib.hnsw_index_get_vector(self._h, label, _ptr(out)))
        return out

    def device(self) -> DeviceIndex:
        h = C.c_void_p()
        _check(_lib.hnsw_index_device(self._h, C.byref(h)))
        return DeviceIndex(h, self)

    def __len__(self) -> int:
        return int(self._info()[0])

    dim = property(lambda self: int(self._info()[1]))
    metric = property(lambda self: METRIC_NAMES[int(self._info()[2])])
    M = property(lambda self: int(self._info()[3]))
    ef_construction = property(lambda self: int(self._info()[4]))
    max_elements = property(lambda self: int(self._info()[5]))
    ef = property(lambda self: int(self._info()[6]), lambda self, v: self.set_ef(v))
    max_level = property(lambda self: int(self._info()[7]))
    read_only = property(lambda self: bool(self._info()[8]))
    memory_usage = property(lambda self: int(self._info()[9]))

    def __repr__(self) -> str:
        # The reference's text (python/bindings.cpp:228-233), under this module's name.
        return (f"<hnswgpu.Index dim={self.dim} metric={self.metric} M={self.M} size={len(self)}"
                + (" read_only" if self.read_only else "") + ">")


class BruteForceIndex:
    """hnsw::BruteForceIndex over the C ABI (exact k-NN on the GPU)."""

    def __init__(self, dim: int, metric: str = "l2"):
        h = C.c_void_p()
        _check(_lib.hnsw_brute_force_create(dim, METRICS[metric], C.byref(h)))
        self._h, self.dim = h, dim

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.hnsw_brute_force_free(h)
            self._h = None

    def add(self, vectors, labels=None) -> None:
        v, _ = _f32(vectors, self.dim, "vectors")
        lab = None if labels is None else np.ascontiguousarray(labels, dtype=np.uint64)
        _check(_lib.hnsw_brute_force_add(self._h, _ptr(v), _ptr(lab), v.shape[0]))

    def search_rows(self, queries, k: int, filter=None):
        q, _ = _f32(queries, self.dim, "queries")
        out = np.empty(q.shape[0] * k, dtype=NEIGHBOR)
        w, size = filter_words(filter)
        _check(_lib.hnsw_brute_force_search_batch(self._h, _ptr(q), q.shape[0], k, _ptr(out), 0, _ptr(w), size))
        return out.reshape(q.shape[0], k)

    def set_option(self, name: str, value: int) -> None:
        _check(_lib.hnsw_brute_force_set_option(self._h, name.encode(), int(value)))

    def get_option(self, name: str) -> int:
        v = C.c_int64()
        _check(_lib.hnsw_brute_force_get_option(self._h, name.encode(), C.byref(v)))
        return v.value
