"""Benchmark: QPS at recall@10 >= 0.95 of the GPU graph search on 1M x 128 L2.

Workload: the integer-valued Gaussian mixture (256 centres U[0,255]^128,
sigma 12, clipped and rounded) at 1,000,000 base vectors, 10,000 queries from
the same mixture; HNSW M=16, ef_construction=200, built by the engine's
builder (the reference's algorithm and file format) with its insertion
searches on the GPU (--host-build: on the host); k=10.

A step is one search_batch of all queries (device buffers, inputs resident in
HBM). For every ef of the sweep: `warmup` untimed steps, then `steps` timed
steps bracketed by synchronisations and timed with HIP events on the stream the
kernels run on. Recall is scored against exact ground truth from the engine's
GPU brute force (exact for this data: every fp32 partial sum is an integer
below 2^24), with 32 queries re-checked in FP64 on the host. QPS at recall 0.95
is interpolated between the two bracketing ef points (log QPS linear in
recall). The roofline figure uses the fast kernel's own event-timed duration
and the algorithmic bytes per query counted by the kernel itself (rows,
layer-0 lists and upper lists read, query in, results out).

A second sweep runs the same steps alternating over two streams (two contexts
of the handle, two output buffers), so that one launch's tail (its last queries
running on a part-empty GPU) overlaps the next launch's start; it is reported
as `pipelined` beside `value`, which stays one launch per step, and its rows
must equal the serial run's.

The CPU baseline is the reference's search_batch (compiled from its sources
into oracle/_ref) on the same index file with all host cores, over the same
queries and ef sweep; its rows are compared with the GPU's row by row. The
reference's add_batch of the same data on the same cores is the build baseline.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "hnsw-engine_amd"))
# The serving section keeps 8 launches in flight on 8 streams, beside the
# index's stream, the search contexts' streams, this script's timing stream and
# the default stream. HIP maps streams onto a pool of hardware queues
# (GPU_MAX_HW_QUEUES) and serialises launches of streams that share one; with a
# pool of 4 slots would share.
# Set before the HIP runtime starts (the first HIP call, after this line).
SERVING_SLOTS = 8
if int(os.environ.get("GPU_MAX_HW_QUEUES", "0") or 0) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

METRIC = "QPS at recall@10≥0.95 (1M×128 L2), GPU FP32 vs BF16+rerank vs CPU; p99 ms"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


MFMA_FP32_PEAK_TFLOPS = 157.3


def ground_truth(X: np.ndarray, Q: np.ndarray, k: int) -> tuple[np.ndarray, float, dict]:
    """Exact top-k labels by the engine's GPU brute force (best k by (distance,
    label)). The data are integers in [0, 255], so every fp32 partial sum is an
    exact integer below 2^24 and the fp32 order equals the FP64 order; a sample
    of queries is re-checked in FP64 on the host. Returns (labels, fraction of
    the sample whose FP64 top-k labels agree, brute-force timings: the MFMA
    prefilter kernel's fp32 matrix rate over 2*nq*n*dim flops and the device
    time of the whole search)."""
    import hnswgpu

    bf = hnswgpu.BruteForceIndex(X.shape[1], "l2")
    bf.add(X)
    bf.set_option("kernel_timing", 1)
    out = bf.search_rows(Q, k)["label"].astype(np.int64)
    gemm, total = [], []
    for _ in range(3):
        again = bf.search_rows(Q, k)["label"].astype(np.int64)
        assert np.array_equal(again, out)
        gemm.append(bf.get_option("last_gemm_ns") / 1e6)
        total.append(bf.get_option("last_total_ns") / 1e6)
    flops = 2.0 * Q.shape[0] * X.shape[0] * X.shape[1]
    g, t = float(np.median(gemm)), float(np.median(total))
    info = {"n": int(X.shape[0]), "nq": int(Q.shape[0]), "k": k, "mfma": bf.get_option("last_mfma"),
            "chunks": bf.get_option("last_chunks"), "scan_requeries": bf.get_option("last_redo"),
            "gemm_ms": round(g, 3), "device_ms": round(t, 3),
            "qps": round(Q.shape[0] / (t / 1e3), 1) if t > 0 else None,
            "mfma_tflops": round(flops / (g / 1e3) / 1e12, 2) if g > 0 else None,
            "mfma_frac": round(flops / (g / 1e3) / 1e12 / MFMA_FP32_PEAK_TFLOPS, 4) if g > 0 else None,
            "mfma_peak_tflops": MFMA_FP32_PEAK_TFLOPS}
    del bf
    agree = 0
    sample = min(32, Q.shape[0])
    for i in range(sample):
        d64 = ((X.astype(np.float64) - Q[i].astype(np.float64)) ** 2).sum(1)
        top = np.lexsort((np.arange(X.shape[0]), d64))[:k]
        agree += int(np.array_equal(top, out[i]))
    return out, agree / sample, info


def tune_report(report: str, fields) -> dict:
    """The autotuner's measurements per ef and dimension: for every value a
    dimension took, the fastest step time (ms) of any shape with that value."""
    out: dict = {}
    for line in report.splitlines():
        if not line.startswith("ef="):
            continue
        kv = dict(item.split("=") for item in line.split())
        ef, ms = kv["ef"], float(kv["ms"])
        for f in fields:
            d = out.setdefault(ef, {}).setdefault(f, {})
            d[kv[f]] = round(min(ms, d.get(kv[f], float("inf"))), 4)
    return out


def measured_traffic(ef: int, args, kernel_ms: float) -> dict:
    """HBM bytes per fast-kernel launch from the committed rocprofv3 counter
    runs (FETCH_SIZE and WRITE_SIZE, scaled by the copy-kernel calibration) on
    this workload, when they cover this ef and shape and still describe the
    kernel that ran: the profile records the kernel's duration in its own run,
    and a duration measured here more than 10 % away from it means the kernel
    changed since (traffic null, with the reason)."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "fast_search_mixture1M_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        rec = t["kernels"][str(ef)]
        if args.n != 1_000_000 or args.nq != t["nq"] or args.k != 10 or args.set:
            raise KeyError("different workload")
    except (OSError, KeyError, ValueError):
        return {"traffic": None}
    prof_ms = rec.get("kernel_ms")
    if prof_ms is None or abs(kernel_ms - prof_ms) > 0.10 * prof_ms:
        return {"traffic": None, "traffic_note": f"profile stale: kernel {kernel_ms:.3f} ms here, "
                                                 f"{prof_ms} ms in {t.get('source', 'the profile')}"}
    return {"traffic": rec["hbm_bytes"], "traffic_source": t.get("source", "profiles/"),
            "traffic_over_algorithmic": round(rec["hbm_bytes"] / rec["algorithmic_bytes"], 3),
            "traffic_profile_kernel_ms": prof_ms}


def serving_bench(ix, Q: np.ndarray, seconds: float, slots: int = 3, max_cost: int = 0, heavy_cost: int = 0) -> dict:
    """The serving loop under the mixed load of a production index: requests
    of log-uniform batch size 1..512, k in {1, 10, 100}, ef in {32..512},
    30 % with a 10 %-density BitsetFilter, half traversing the BF16 copy, 1 %
    of the base soft-deleted. Closed loop (all requests at once: saturated
    throughput), then open loop (Poisson arrivals at half that request rate):
    QPS beside p50 / p99 request latency (submit to completion)."""
    import queue
    import threading

    import hnswgpu

    rng = np.random.default_rng(77)
    n = len(ix)
    for label in rng.choice(n, n // 100, replace=False):
        ix.mark_deleted(int(label))
    dev = ix.device()
    dev.set_option("storage", 0)
    fw = hnswgpu.filter_words(rng.random(n) < 0.10)  # one filter shared by the filtered requests

    def make(count):
        out = []
        for _ in range(count):
            nq = int(np.exp(rng.uniform(0.0, np.log(512.0))))
            s = int(rng.integers(0, Q.shape[0] - nq))
            out.append((Q[s:s + nq], int(rng.choice([1, 10, 100])), int(rng.choice([32, 64, 128, 256, 512])),
                        fw if rng.random() < 0.3 else None, 1 if rng.random() < 0.5 else 0))
        return out

    def summary(st, wall, reqs):
        nq = sum(r[0].shape[0] for r in reqs)
        return {"requests": len(reqs), "queries": nq, "wall_s": round(wall, 3), "qps": round(nq / wall, 1),
                "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3), "max_ms": round(st["max_ms"], 3),
                "launches": int(st["launches"]), "mean_queries_per_launch": round(st["mean_batch"], 1),
                "graph_launches": int(st["graph_launches"])}

    srv = hnswgpu.Server(dev, slots=slots, max_batch=4096, storages=(1,), max_cost=max_cost, heavy_cost=heavy_cost)
    # Warm-up: every slot captures one graph per (storage, ef class, kernel)
    # on first use, the kernel being the mixed one when a launch holds a
    # filtered request and the deletions side-list one when it holds none.
    # One request of a full launch per slot occupies every slot at once, for
    # every combination; then a random mix.
    for st in (0, 1):
        for ef in (32, 128, 256, 512):
            for f in (None, fw):
                ts = [srv.submit(Q[:4096], 10, ef, f, storage=st) for _ in range(slots)]
                for t in ts:
                    srv.wait(t)
    warm = make(50)
    for t in [srv.submit(q, k, ef, f, storage=st) for q, k, ef, f, st in warm]:
        srv.wait(t)
    reqs = make(3000)
    srv.reset_stats()
    t0 = time.perf_counter()
    tickets = [srv.submit(q, k, ef, f, storage=st) for q, k, ef, f, st in reqs]
    for t in tickets:
        srv.wait(t)
    closed = summary(srv.stats(), time.perf_counter() - t0, reqs)
    log(f"serving closed loop {closed}")

    # A request is heavy when its queries cost at least the server's heavy_cost
    # (max(ef, k) / filter density >= 1,024: the filtered ones at ef >= 128);
    # heavy requests form launches of their own.
    def heavy(r):
        return (max(r[2], r[1]) / (0.10 if r[3] is not None else 1.0)) >= 1024

    def by_class(reqs, lat):
        out = {}
        for name, sel in (("cheap", lambda r: not heavy(r)), ("heavy", heavy)):
            ls = np.array([l for r, l in zip(reqs, lat) if sel(r)])
            if ls.size:
                out[name] = {"requests": int(ls.size), "p50_ms": round(float(np.percentile(ls, 50)), 3),
                             "p99_ms": round(float(np.percentile(ls, 99)), 3)}
        return out

    def open_loop(rate, seconds_):
        reqs = make(max(100, int(rate * seconds_)))
        srv.reset_stats()
        pending: "queue.Queue[int | None]" = queue.Queue()
        lat = []

        def waiter():
            while (t := pending.get()) is not None:
                lat.append(srv.wait_latency(t))

        th = threading.Thread(target=waiter)
        th.start()
        t0 = time.perf_counter()
        due = t0
        for q, k, ef, f, st in reqs:
            due += rng.exponential(1.0 / rate)
            while (now := time.perf_counter()) < due:
                if due - now > 2e-4:
                    time.sleep(due - now - 1e-4)
            pending.put(srv.submit(q, k, ef, f, storage=st))
        pending.put(None)
        th.join()
        out = summary(srv.stats(), time.perf_counter() - t0, reqs)
        out["offered_requests_per_s"] = round(rate, 1)
        out["by_class"] = by_class(reqs, lat)
        return out

    saturated = closed["requests"] / closed["wall_s"]
    opened = {}
    for frac in (0.25, 0.5, 0.75):
        opened[str(frac)] = open_loop(frac * saturated, seconds)
        log(f"serving open loop at {frac:.2f} of the saturated request rate {opened[str(frac)]}")
    srv.close()
    return {"mix": "batch 1..512 log-uniform, k {1,10,100}, ef {32..512}, 30% filtered (10% density), "
                   f"50% bf16 traversal, 1% deleted; {slots} slots, max 4096 queries and 2^20 cost units "
                   "(ef / filter density) per launch, graphs",
            "closed_loop": closed,
            "open_loop_poisson": opened["0.5"],
            "open_loop_by_load": opened}


def recall_at(labels: np.ndarray, truth: np.ndarray, k: int) -> float:
    hits = 0
    for a, b in zip(labels[:, :k], truth[:, :k]):
        hits += len(set(a.tolist()) & set(b.tolist()))
    return hits / float(truth.shape[0] * k)


def qps_at(points, target=0.95):
    import workloads

    return workloads.qps_at_recall(points, target)


def cpu_baseline(index_path: str, Q: np.ndarray, truth: np.ndarray, k: int, efs, budget_s: float, X=None,
                 lat_ef: int = 128):
    """The reference's search_batch on this host's cores (oracle/_ref), its
    single-query latency on one thread at lat_ef, and, when X is given, the
    reference's add_batch of X on the same cores (the build baseline; the time
    includes writing the file)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    from oracle import REF_BENCH_SO, REF_SO, Reference

    cores = len(os.sched_getaffinity(0))
    # The reference as its bench preset builds it (-O3 -march=native, pinned to the AVX-512 level
    # x86-64-v4), else the portable checker build (-O3, per-file ISA flags).
    so, flags = ((REF_BENCH_SO, "-O3 -march=x86-64-v4") if REF_BENCH_SO.exists()
                 else (REF_SO, "-O3, -mavx2 -mfma / -mavx512f -mfma on the kernel files only"))
    if not so.exists():
        return None
    ref = Reference(so)
    ref_build_s = None
    if X is not None:
        path = os.path.join(tempfile.gettempdir(), "hnsw_bench_reference_build.idx")
        t = time.perf_counter()
        ref.build_save(path, X, metric=0, M=16, efc=200, threads=cores)
        ref_build_s = time.perf_counter() - t
        os.remove(path)
        log(f"reference add_batch of {X.shape[0]} x {X.shape[1]} on {cores} threads (+ save): {ref_build_s:.1f}s")
    idx = ref.load(index_path, mmap=True)
    # Bounded sample: size it from a short probe at the largest ef.
    probe = Q[:200]
    t = time.perf_counter()
    idx.search(probe, k, max(efs), threads=cores)
    per_q = (time.perf_counter() - t) / probe.shape[0]
    nq = int(min(Q.shape[0], max(500, budget_s / max(per_q * len(efs) * 2, 1e-9))))
    sample = Q[:nq]
    points, rows_by_ef = [], {}
    for ef in efs:
        idx.search(sample[:200], k, ef, threads=cores)  # warm
        t = time.perf_counter()
        out = idx.search(sample, k, ef, threads=cores)
        dt = time.perf_counter() - t
        rows_by_ef[ef] = out
        rec = recall_at(out["label"].astype(np.int64), truth[:nq], k)
        points.append((ef, rec, nq / dt))
        log(f"cpu reference ef={ef} recall={rec:.4f} qps={nq / dt:,.0f} ({cores} threads)")
    # One query per call on one thread, as the GPU's single-query latency is measured.
    lat = []
    for i in range(220):
        t = time.perf_counter()
        idx.search(Q[i:i + 1], k, lat_ef, threads=1)
        lat.append((time.perf_counter() - t) * 1e3)
    lat = np.array(lat[20:])
    log(f"cpu reference single query ef={lat_ef}: p50 {np.percentile(lat, 50):.4f} ms, p99 {np.percentile(lat, 99):.4f} ms")
    return {"value": qps_at(points), "unit": "queries/s", "cores": cores, "kind": "reference",
            "isa": ref.active_isa(), "build": f"{so.name}: {flags}", "reference_build_s": round(ref_build_s, 2) if ref_build_s else None,
            "sample": f"first {nq} of the {Q.shape[0]} queries, ef sweep {list(efs)}, search_batch on {cores} threads",
            "points": [{"ef": e, "recall": round(r, 5), "qps": round(q, 1)} for e, r, q in points],
            "single_query_1thread_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                                        "p99": round(float(np.percentile(lat, 99)), 4), "at_ef": lat_ef,
                                        "queries": int(lat.size)},
            "_rows": rows_by_ef}


def stored_distance_error(dev, X: np.ndarray, Q: np.ndarray, storage: int, metric: str, rows: int = 2000,
                          queries: int = 100) -> dict:
    """Traversal distance of the stored copy (bf16 rows, or e4m3 codes x scale) against the FP64
    distance of the original vectors, on random (query, row) pairs: relative error quantiles.
    The stored values are read back from the device; the arithmetic is done in FP64, so this is
    the error of the stored copy (fp32 accumulation adds ~1e-7 relative)."""
    rng = np.random.default_rng(77)
    ids = rng.choice(X.shape[0], rows, replace=False)
    P = int(np.ceil(X.shape[1] / 16) * 16)
    if storage == 1:
        vals = np.stack([dev.get_row(int(i), 1, P)[0] for i in ids]).astype(np.uint32)
        xs = (vals << 16).view(np.float32).astype(np.float64)[:, :X.shape[1]]
    else:
        tab = np.zeros(256)
        for code in range(256):
            s_, e, m = code >> 7, (code >> 3) & 0xF, code & 7
            v = (m / 8.0) * 2.0 ** -6 if e == 0 else (1 + m / 8.0) * 2.0 ** (e - 7)
            tab[code] = np.nan if (e == 0xF and m == 7) else (-v if s_ else v)
        got = [dev.get_row(int(i), 2, P) for i in ids]
        xs = np.stack([tab[r] * sc for r, sc, _ in got])[:, :X.shape[1]]
    x64 = X[ids].astype(np.float64)
    q64 = Q[:queries].astype(np.float64)
    if metric == "cosine":
        x64 = x64 / np.linalg.norm(x64, axis=1, keepdims=True)
        q64 = q64 / np.linalg.norm(q64, axis=1, keepdims=True)
        qr = q64
        if storage == 1:  # the query is rounded to bf16 as well
            u = q64.astype(np.float32).view(np.uint32).astype(np.uint64)
            qr = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)
        exact = 1.0 - q64 @ x64.T
        approx = 1.0 - qr @ xs.T
    else:
        exact = (q64 ** 2).sum(1)[:, None] + (x64 ** 2).sum(1)[None, :] - 2.0 * q64 @ x64.T
        approx = (q64 ** 2).sum(1)[:, None] + (xs ** 2).sum(1)[None, :] - 2.0 * q64 @ xs.T
    rel = np.abs(approx - exact) / np.maximum(np.abs(exact), 1e-12)
    return {"pairs": int(rel.size), "median_rel": float(f"{np.median(rel):.3e}"),
            "p99_rel": float(f"{np.quantile(rel, 0.99):.3e}"), "max_rel": float(f"{rel.max():.3e}")}


def reduced_precision_run(args, which: str) -> dict:
    """A reduced-precision configuration at its own size: 1M x 100 cosine with bf16 rows (k 10 and
    100; `which` = cosine100) or 200k x 960 L2 with fp8 e4m3 rows + per-vector scale (k 10;
    sparse960). It sweeps ef for the fp32 traversal and the reduced one (exact fp32 re-rank of the
    final candidates) and reports QPS at recall 0.95 for both and their ratio, the recall at every
    ef, the stored copy's distance error against FP64, and the ground truth's agreement with FP64
    (GPU brute force, checked in FP64 on a sample)."""
    import hnswgpu
    import workloads

    if which == "cosine100":
        n, nq, metric, storage, ks, efs = 1_000_000, 10_000, "cosine", 1, (10, 100), (32, 64, 128, 256, 512)
        X, Q = workloads.cosine_100(n, seed=3), workloads.cosine_100(nq, seed=1003)
        name = "1M x 100 cosine (1,000-centre mixture, log-normal norms), bf16 rows + fp32 re-rank"
    else:
        n, nq, metric, storage, ks, efs = 200_000, 5_000, "l2", 2, (10,), (64, 128, 256)
        X, Q = workloads.sparse_960(n, seed=4), workloads.sparse_960(nq, seed=1004)
        name = "200k x 960 L2 (|N(0,1)| x 0.1, 10 % of dims zeroed), fp8 e4m3 rows + scale + fp32 re-rank"
    dim = X.shape[1]
    threads = args.build_threads or len(os.sched_getaffinity(0))
    t0 = time.perf_counter()
    ix = hnswgpu.Index(dim, metric, M=16, ef_construction=200)
    ix.add(X, num_threads=threads, device=not args.host_build)
    log(f"built {n} x {dim} {metric} ({'host' if args.host_build else 'device'} probe, {threads} host threads) "
        f"in {time.perf_counter() - t0:.1f}s (max level {ix.max_level})")
    kmax = max(ks)
    bf = hnswgpu.BruteForceIndex(dim, metric)
    bf.add(X)
    truth = bf.search_rows(Q, kmax)["label"].astype(np.int64)
    del bf
    agree, sample = 0, 16
    x64 = X.astype(np.float64)
    if metric == "cosine":
        x64 /= np.linalg.norm(x64, axis=1, keepdims=True)
    for i in range(sample):
        q = Q[i].astype(np.float64)
        d = (1.0 - x64 @ (q / np.linalg.norm(q))) if metric == "cosine" else ((x64 - q) ** 2).sum(1)
        agree += int(np.array_equal(np.lexsort((np.arange(n), d))[:kmax], truth[i]))
    del x64
    log(f"ground truth: GPU brute force, FP64 top-{kmax} agreement on {sample} queries {agree / sample:.3f}")
    dev = ix.device()
    dev.autotune(Q, ks[0], [e for e in efs if e <= 512])
    stream = hnswgpu.Stream()
    d_q = hnswgpu.DeviceBuffer.of(Q)
    d_out = hnswgpu.DeviceBuffer(nq * kmax * hnswgpu.NEIGHBOR.itemsize)
    result = {"workload": name, "n": n, "dim": dim, "nq": nq, "metric": metric,
              "storage": {1: "bf16", 2: "fp8_e4m3_scale"}[storage], "fp64_truth_agreement": agree / sample, "k": {}}
    for k in ks:
        host_rows = np.empty(nq * k, hnswgpu.NEIGHBOR)
        per = {}
        for st in (0, storage):
            dev.set_option("storage", st)
            pts = []
            for ef in efs:
                if ef < k:
                    continue
                for _ in range(args.warmup):
                    dev.search_device(d_q.ptr, nq, k, ef, d_out.ptr, stream=stream.handle)
                stream.synchronize()
                start, end = hnswgpu.Event(), hnswgpu.Event()
                start.record(stream)
                for _ in range(args.steps):
                    dev.search_device(d_q.ptr, nq, k, ef, d_out.ptr, stream=stream.handle)
                end.record(stream)
                stream.synchronize()
                ms = start.elapsed_ms(end) / args.steps
                rows = d_out.download(host_rows).reshape(nq, k)
                rec = recall_at(rows["label"].astype(np.int64), truth, k)
                pts.append((ef, rec, nq / (ms / 1e3)))
                log(f"{which} k={k} storage={st} ef={ef} recall={rec:.4f} qps={nq / (ms / 1e3):,.0f} step={ms:.3f}ms")
            v = qps_at(pts)
            per["fp32" if st == 0 else "reduced"] = {
                "qps_at_recall_0.95": round(v, 1) if math.isfinite(v) else None,
                "points": [{"ef": e, "recall": round(r, 5), "qps": round(q, 1)} for e, r, q in pts]}
        dev.set_option("storage", 0)
        a, b = per["fp32"]["qps_at_recall_0.95"], per["reduced"]["qps_at_recall_0.95"]
        per["reduced_over_fp32"] = round(b / a, 3) if a and b else None
        # The same ef on both copies, with each side's recall: also defined where neither
        # reaches the recall target.
        per["by_ef"] = [{"ef": pa["ef"], "reduced_over_fp32": round(pb["qps"] / pa["qps"], 3),
                         "recall_fp32": pa["recall"], "recall_reduced": pb["recall"]}
                        for pa, pb in zip(per["fp32"]["points"], per["reduced"]["points"]) if pa["ef"] == pb["ef"]]
        result["k"][str(k)] = per
    dev.set_option("storage", storage)
    result["stored_distance_error_vs_fp64"] = stored_distance_error(dev, X, Q, storage, metric)
    dev.set_option("storage", 0)
    return result


def precision_workload(args) -> int:
    """One JSON line for one reduced-precision configuration (--workload)."""
    import hnswgpu

    log(f"device {hnswgpu.device_info()}")
    print(json.dumps(reduced_precision_run(args, args.workload)), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--nq", type=int, default=10_000)
    ap.add_argument("--k", type=int, default=10)
    ap.add_argument("--efs", default="16,32,64,128,256")
    ap.add_argument("--build-threads", type=int, default=0)
    ap.add_argument("--cpu-budget-s", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--fp32-only", action="store_true", help="skip the bf16 / fp8 traversal sweeps")
    ap.add_argument("--index", default="", help="reuse/save the built index at this path")
    ap.add_argument("--profile", action="store_true",
                    help="profiling run: no ground truth, latency or CPU baseline; one line with the timings")
    ap.add_argument("--deleted", type=float, default=0.0, metavar="FRAC",
                    help="profiling run only: soft-delete this fraction of the base (seeded) before the sweep")
    ap.add_argument("--filter", type=float, default=0.0, metavar="DENSITY",
                    help="profiling run only: search with a seeded label filter allowing this fraction of labels")
    ap.add_argument("--set", action="append", default=[], metavar="NAME=VALUE",
                    help="device-layer option for every search (e.g. hash_log2=12)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device (counter calibration) and exit")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-loop section")
    ap.add_argument("--no-pipelined", action="store_true", help="skip the two-stream pipelined sweep")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the 1M x 100 cosine bf16 and 200k x 960 fp8 configurations")
    ap.add_argument("--no-autotune", action="store_true", help="default launch shapes instead of the autotuner's")
    ap.add_argument("--host-build", action="store_true",
                    help="build the index with the host probe only (default: insertion searches on the GPU)")
    ap.add_argument("--serving-seconds", type=float, default=3.0)
    ap.add_argument("--serving-slots", type=int, default=SERVING_SLOTS)
    ap.add_argument("--serving-max-cost", type=int, default=0, help="cost units per launch (0: the server's default)")
    ap.add_argument("--serving-heavy-cost", type=int, default=0,
                    help="query cost from which requests form launches of their own (0: the server's default)")
    ap.add_argument("--serving-only", action="store_true", help="only the serving section, one line")
    ap.add_argument("--profile-brute", action="store_true",
                    help="profiling run of the brute force only (n x 128 data, nq queries, k), one line")
    ap.add_argument("--workload", default="mixture1m", choices=["mixture1m", "cosine100", "sparse960"],
                    help="mixture1m: the headline line; cosine100 / sparse960: the reduced-precision "
                         "workloads (1M x 100 cosine with bf16, 200k x 960 L2 with fp8 e4m3), one line each")
    args = ap.parse_args()
    if args.calibrate_copy:
        import hnswgpu

        nbytes = args.calibrate_copy * (1 << 20)
        src, dst = hnswgpu.DeviceBuffer(nbytes), hnswgpu.DeviceBuffer(nbytes)
        src.upload(np.arange(nbytes // 4, dtype=np.float32))
        for _ in range(3):
            hnswgpu.copy_device(dst, src, nbytes)
        print(json.dumps({"copy_bytes": nbytes, "launches": 3}), flush=True)
        return 0
    if args.workload != "mixture1m" and args.gpus == 1:
        return precision_workload(args)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the index is replicated "
              "per GPU with no exchange, so N GPUs serve N independent query streams).", file=sys.stderr)
        return 2

    import hnswgpu
    import workloads

    log(f"device {hnswgpu.device_info()}")
    k = args.k
    efs = [int(e) for e in args.efs.split(",")]

    t0 = time.perf_counter()
    X = workloads.mixture_128(args.n, seed=5)
    Q = workloads.mixture_128(args.nq, seed=1005)
    log(f"data {X.shape} + {Q.shape} in {time.perf_counter() - t0:.1f}s")
    if args.profile_brute:
        bf = hnswgpu.BruteForceIndex(X.shape[1], "l2")
        bf.add(X)
        bf.set_option("kernel_timing", 1)
        runs = []
        for _ in range(3):
            bf.search_rows(Q, k)
            runs.append({"gemm_ms": bf.get_option("last_gemm_ns") / 1e6, "device_ms": bf.get_option("last_total_ns") / 1e6,
                         "scan_requeries": bf.get_option("last_redo"), "chunks": bf.get_option("last_chunks")})
        print(json.dumps({"profile_brute": True, "n": args.n, "nq": args.nq, "k": k, "runs": runs}), flush=True)
        return 0

    index_path = args.index or os.path.join(tempfile.gettempdir(), f"hnsw_bench_{args.n}.idx")
    build = None
    if args.index and os.path.exists(index_path):
        ix = hnswgpu.Index.load(index_path)
        log(f"loaded {index_path}")
    else:
        threads = args.build_threads or len(os.sched_getaffinity(0))
        t0 = time.perf_counter()
        ix = hnswgpu.Index(128, "l2", M=16, ef_construction=200)
        stats = ix.add(X, num_threads=threads, device=not args.host_build)
        build_s = time.perf_counter() - t0
        build = {"builder": "host" if args.host_build else "device", "seconds": round(build_s, 2),
                 "host_threads": threads, "max_level": ix.max_level}
        if stats:
            build.update({k: (round(v, 1) if isinstance(v, float) else v) for k, v in stats.items()})
        log(f"built {args.n} x 128 ({build['builder']} probe, {threads} host threads) in {build_s:.1f}s: {build}")
        ix.save(index_path)

    if args.serving_only:
        print(json.dumps({"serving": serving_bench(ix, Q, args.serving_seconds, args.serving_slots,
                                                   args.serving_max_cost, args.serving_heavy_cost)}), flush=True)
        return 0

    if args.deleted > 0:
        if not args.profile:
            raise SystemExit("--deleted is for profiling runs (--profile)")
        for lab in np.random.default_rng(77).permutation(len(ix))[: int(len(ix) * args.deleted)]:
            ix.mark_deleted(int(lab))
        log(f"soft-deleted {int(len(ix) * args.deleted)} of {len(ix)} labels")

    t0 = time.perf_counter()
    truth, truth_check, brute = (None, None, None) if args.profile else ground_truth(X, Q, k)
    log(f"ground truth in {time.perf_counter() - t0:.1f}s (FP64 check agreement {truth_check}); brute force {brute}")

    dev = ix.device()
    tuned = None
    if args.set or args.no_autotune or args.profile:  # a profile times the default launch shapes
        for kv in args.set:
            name, value = kv.split("=")
            dev.set_option(name, int(value))
    else:
        # The autotuner picks the launch shape per ef class on batches of the size timed below.
        t0 = time.perf_counter()
        report = dev.autotune(Q, k, [min(e, 512) for e in efs])
        fields = ("waves_per_cu", "hash_log2", "generic_gather", "row_trip", "min_class")
        tuned = {str(c): {f: dev.get_option(f"tuned_{f}_{c}") for f in fields}
                 for c in (64, 128, 256, 512) if dev.get_option(f"tuned_waves_per_cu_{c}") >= 0}
        tuned = {"chosen": tuned, "by_dimension": tune_report(report, fields),
                 "seconds": round(time.perf_counter() - t0, 2)}
        log(f"autotune in {time.perf_counter() - t0:.1f}s: {tuned}")
    stream = hnswgpu.Stream()
    d_q = hnswgpu.DeviceBuffer.of(Q)
    d_out = hnswgpu.DeviceBuffer(args.nq * k * hnswgpu.NEIGHBOR.itemsize)
    host_rows = np.empty(args.nq * k, hnswgpu.NEIGHBOR)
    d_filter, filter_size = None, 0
    if args.filter > 0:
        if not args.profile:
            raise SystemExit("--filter is for profiling runs (--profile)")
        fw, filter_size = hnswgpu.filter_words(np.random.default_rng(78).random(len(ix)) < args.filter)
        d_filter = hnswgpu.DeviceBuffer.of(fw)
        log(f"label filter allowing {args.filter:.0%} of {len(ix)} labels")

    def step(ef):
        dev.search_device(d_q.ptr, args.nq, k, ef, d_out.ptr, stream=stream.handle,
                          d_filter=d_filter.ptr if d_filter is not None else 0, filter_size=filter_size)

    points, per_ef, gpu_rows = [], {}, {}
    for ef in efs:
        for _ in range(args.warmup):
            step(ef)
        stream.synchronize()
        start, end = hnswgpu.Event(), hnswgpu.Event()
        start.record(stream)
        for _ in range(args.steps):
            step(ef)
        end.record(stream)
        stream.synchronize()
        ms = start.elapsed_ms(end) / args.steps
        rows = d_out.download(host_rows).reshape(args.nq, k)
        gpu_rows[ef] = rows.copy()
        rec = float("nan") if truth is None else recall_at(rows["label"].astype(np.int64), truth, k)
        qps = args.nq / (ms / 1e3)
        # Fast-kernel time per launch (events on the same stream) and work counts.
        dev.set_option("kernel_timing", 1)
        kns = []
        for _ in range(max(3, min(args.steps, 10))):
            step(ef)
            kns.append(dev.get_option("last_kernel_ns"))
        dev.set_option("kernel_timing", 0)
        dev.set_option("collect_stats", 1)
        step(ef)
        stats = dev.read_stats(args.nq).astype(np.float64)
        dev.set_option("collect_stats", 0)
        overflow = dev.get_option("last_overflow")
        spilled = dev.get_option("last_hbm_visited")
        handoff_c = dev.get_option("last_handoff_candidates")
        n_dist, n_exp0, n_expu = stats.mean(axis=0)
        bytes_q = n_dist * 128 * 4 + n_exp0 * 4 * (1 + 32) + n_expu * 4 * (1 + 16) + 128 * 4 + k * 16
        kernel_s = float(np.mean(kns)) / 1e9
        gbs = bytes_q * args.nq / kernel_s / 1e9
        per_ef[ef] = {"ef": ef, "recall": round(rec, 5), "qps": round(qps, 1), "ms_per_step": round(ms, 4),
                      "kernel_ms": round(kernel_s * 1e3, 4), "bytes_per_query": round(bytes_q, 1),
                      "dist_per_query": round(n_dist, 1), "hbm_gbs": round(gbs, 1),
                      "general_kernel_queries": int(overflow), "hbm_visited_queries": int(spilled),
                      "candidate_handoffs": int(handoff_c)}
        points.append((ef, rec, qps))
        log(f"gpu ef={ef} recall={rec:.4f} qps={qps:,.0f} step={ms:.3f}ms kernel={kernel_s * 1e3:.3f}ms "
            f"bytes/q={bytes_q:,.0f} achieved={gbs:,.0f} GB/s dist/q={n_dist:.0f} general={overflow} "
            f"hbm-visited={spilled}")

    # Pipelined: the same steps alternating over two streams (two contexts of the handle, two
    # output buffers), so that a launch's tail, when its last queries run on a part-empty GPU,
    # overlaps the next launch's start. Each step is still one complete search of the batch.
    pipelined = None
    if not args.profile and not args.no_pipelined:
        stream2 = hnswgpu.Stream()
        d_out2 = hnswgpu.DeviceBuffer(args.nq * k * hnswgpu.NEIGHBOR.itemsize)
        pp = []
        for ef in efs:
            def step2(i, ef=ef):
                st, out = (stream, d_out) if i % 2 == 0 else (stream2, d_out2)
                dev.search_device(d_q.ptr, args.nq, k, ef, out.ptr, stream=st.handle)
            for i in range(max(2, args.warmup)):
                step2(i)
            stream.synchronize()
            stream2.synchronize()
            t0 = time.perf_counter()
            for i in range(args.steps):
                step2(i)
            stream.synchronize()
            stream2.synchronize()
            dt = (time.perf_counter() - t0) / args.steps
            rows2 = d_out2.download(host_rows).reshape(args.nq, k)
            same = bool(np.array_equal(rows2["label"], gpu_rows[ef]["label"]) and
                        np.array_equal(rows2["distance"].view(np.uint32), gpu_rows[ef]["distance"].view(np.uint32)))
            rec = per_ef[ef]["recall"]
            pp.append((ef, rec, args.nq / dt))
            log(f"gpu pipelined ef={ef} qps={args.nq / dt:,.0f} step={dt * 1e3:.3f}ms rows equal the serial run: {same}")
            if not same:
                raise SystemExit("pipelined rows differ from the serial run")
        v2 = qps_at(pp)
        pipelined = {"streams": 2, "value": round(v2, 1) if math.isfinite(v2) else None,
                     "over_serial": round(v2 / qps_at(points), 3) if math.isfinite(v2) else None,
                     # Algorithmic bytes of a step over the pipelined step time, against 8 TB/s.
                     "points": [{"ef": e, "recall": r, "qps": round(q, 1),
                                 "hbm_frac": round(per_ef[e]["bytes_per_query"] * q / 1e9 / HBM_PEAK_GBS, 4)}
                                for e, r, q in pp]}
        del d_out2, stream2

    if args.profile:
        print(json.dumps({"profile": True, "points": list(per_ef.values())}), flush=True)
        return 0
    value = qps_at(points)

    # Reduced-precision traversal (bf16; fp8 e4m3 + scale) with exact fp32 re-rank.
    reduced = {}
    for storage, name in ((1, "bf16_rerank"), (2, "fp8_rerank")):
        if args.fp32_only:
            break
        dev.set_option("storage", storage)
        rp = []
        for ef in efs:
            for _ in range(args.warmup):
                step(ef)
            stream.synchronize()
            start, end = hnswgpu.Event(), hnswgpu.Event()
            start.record(stream)
            for _ in range(args.steps):
                step(ef)
            end.record(stream)
            stream.synchronize()
            ms = start.elapsed_ms(end) / args.steps
            rows = d_out.download(host_rows).reshape(args.nq, k)
            rec = recall_at(rows["label"].astype(np.int64), truth, k)
            rp.append((ef, rec, args.nq / (ms / 1e3)))
            log(f"gpu {name} ef={ef} recall={rec:.4f} qps={args.nq / (ms / 1e3):,.0f} step={ms:.3f}ms")
        dev.set_option("storage", 0)
        v = qps_at(rp)
        reduced[name] = {"value": round(v, 1) if math.isfinite(v) else None,
                         "over_fp32": round(v / value, 3) if math.isfinite(v) and math.isfinite(value) else None,
                         "points": [{"ef": e, "recall": round(r, 5), "qps": round(q, 1)} for e, r, q in rp]}

    # Roofline at the first ef that reaches the recall target.
    op = next((per_ef[e] for e, r, _ in points if r >= 0.95), per_ef[efs[-1]])

    # Single-query latency through the host API (copy in, search, copy out).
    lat = []
    q1 = Q[:1]
    for i in range(220):
        t = time.perf_counter()
        ix.search_rows(Q[i % args.nq:i % args.nq + 1], k, op["ef"])
        lat.append((time.perf_counter() - t) * 1e3)
    lat = np.array(lat[20:])
    del q1
    lat_default = None
    if tuned is not None:
        # The same with the default launch shapes (the tuner's shapes off).
        dev.set_option("use_tuned", 0)
        ld = []
        for i in range(220):
            t = time.perf_counter()
            ix.search_rows(Q[i % args.nq:i % args.nq + 1], k, op["ef"])
            ld.append((time.perf_counter() - t) * 1e3)
        dev.set_option("use_tuned", 1)
        lat_default = np.array(ld[20:])

    cpu, parity = None, None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(index_path, Q, truth, k, efs, args.cpu_budget_s, X=X if build is not None else None,
                           lat_ef=op["ef"])
    if cpu is not None:
        # Row parity with the reference's own search_batch on the same file and queries: a row
        # mismatches if any label or distance bit differs (the GPU computes in the AVX-512 order).
        parity = {"against": f"reference search_batch ({cpu['isa']} kernels)", "mismatching_rows": {}}
        for ef, ref_rows in cpu.pop("_rows").items():
            n = ref_rows.shape[0]
            g = gpu_rows[ef][:n]
            bad = ~((g["label"] == ref_rows["label"]).all(1) &
                    (g["distance"].view(np.uint32) == ref_rows["distance"].view(np.uint32)).all(1))
            parity["mismatching_rows"][str(ef)] = int(bad.sum())
            parity["rows_compared"] = int(n)
        log(f"parity with the reference: {parity}")

    # Last: it soft-deletes 1 % of the base.
    serving = None if args.no_serving else serving_bench(ix, Q, args.serving_seconds, args.serving_slots,
                                                         args.serving_max_cost, args.serving_heavy_cost)
    del ix, dev
    precision = {}
    if not args.no_precision:
        for which, key in (("cosine100", "cosine100_bf16"), ("sparse960", "sparse960_fp8")):
            t0 = time.perf_counter()
            precision[key] = reduced_precision_run(args, which)
            log(f"{which} in {time.perf_counter() - t0:.1f}s")

    line = {
        "metric": METRIC,
        "value": round(value, 1) if math.isfinite(value) else None,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(args.nq / value * 1e3, 4) if math.isfinite(value) else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "ground_truth": {"method": "GPU brute force (fp32 exact on integer data)", "fp64_sample_agreement": truth_check},
        "data": "synthetic",
        "config": {"workload": f"mixture_128 {args.n}x128 L2 (256 centres, sigma 12, integer-valued), "
                               f"{args.nq} queries, M=16, ef_construction=200, k={k}, ef sweep {efs}",
                   "n": args.n, "nq": args.nq, "k": k, "efs": efs, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": op["hbm_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(op["hbm_gbs"] / HBM_PEAK_GBS, 4), **measured_traffic(op["ef"], args, op["kernel_ms"]),
                     "kernel": "fast_search", "at_ef": op["ef"], "kernel_ms": op["kernel_ms"],
                     "bytes_per_query": op["bytes_per_query"]},
        "cpu_baseline": cpu,
        "pipelined": pipelined,
        "parity": parity,
        "build": build,
        "brute_force": brute,
        "autotune": tuned,
        "serving": serving,
        **reduced,
        **precision,
        "latency_ms": {"single_query_p50": round(float(np.percentile(lat, 50)), 4),
                       "single_query_p99": round(float(np.percentile(lat, 99)), 4),
                       "single_query_p50_default_shapes": (round(float(np.percentile(lat_default, 50)), 4)
                                                           if lat_default is not None else None),
                       "at_ef": op["ef"]},
        "points": list(per_ef.values()),
        "overflow_queries": int(overflow),
    }
    if cpu and cpu.get("value") and math.isfinite(cpu["value"]) and line["value"]:
        line["gpu_over_cpu"] = round(line["value"] / cpu["value"], 2)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
