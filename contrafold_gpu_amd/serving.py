"""Continuous-batching fold server (Python face of cf_server_* in the C ABI).

    srv = Server(engine)
    rid = srv.submit("GGGGAAAACCCC", kind="mea", gamma=6.0)
    while not done: for rid in srv.poll(): res = srv.result(rid)

Requests are admitted while their O(L^2) device table bytes fit the HBM
budget, batched by kind on a pool of HIP streams, and retired as their
batches finish; batches launched while others run overlap with them.  A
request's sampling stream depends only on its sequence, mask and seed, so its
result is the same alone and in any mix.
"""
import ctypes
import time
import zlib

import numpy as np

from ._lib import CF_BPP, CF_LOGZ, CF_MEA, CF_SAMPLE, check

KINDS = {"mea": CF_MEA, "bpp": CF_BPP, "sample": CF_SAMPLE, "logz": CF_LOGZ}
_c = ctypes


def request_key(seq, mask=None, seed=0):
    """64-bit sampling-stream key of a request (content and seed)."""
    b = seq.encode("ascii", "replace")
    if mask is not None:
        b += np.asarray(mask, np.uint8).tobytes()
    h = (zlib.crc32(b) << 32) | zlib.adler32(b)
    return (h ^ ((int(seed) * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF)) & 0xFFFFFFFFFFFFFFFF


class Server:
    """Continuous-batching fold server over one engine (cf_server_*).

    Each of the n_streams slots runs its batches on its own HIP stream.  HIP
    maps streams onto GPU_MAX_HW_QUEUES hardware queues (4 by default) and a
    queue runs its work in order, so set GPU_MAX_HW_QUEUES >= n_streams in
    the environment before the process first uses the GPU; otherwise short
    folds wait behind long ones that share their queue.
    """

    def __init__(self, engine, hbm_budget=64 << 30, max_batch=512, n_streams=16):
        self.engine = engine
        self.lib = engine.lib
        h = _c.c_void_p()
        # the server drives the engine's context: its calls hold the engine's lock
        self.lock = engine.lock
        with self.lock:
            check(self.lib.cf_server_create(engine.ctx, int(hbm_budget), int(max_batch), int(n_streams),
                                            _c.byref(h)), engine.ctx, "cf_server_create")
        self.h = h
        self.meta = {}

    def close(self):
        if getattr(self, "h", None):
            with self.lock:
                self.lib.cf_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, seq, kind="mea", gamma=6.0, n_samples=10, mask=None, seed=0):
        k = KINDS[kind]
        L = len(seq)
        m = None
        if mask is not None:
            arr = np.asarray(mask).reshape(-1)
            if arr.size < L:
                raise ValueError(f"mask has {arr.size} entries for a sequence of length {L}")
            m = np.ascontiguousarray(arr[:L] != 0, np.uint8)
        key = request_key(seq, m, seed)
        with self.lock:
            rid = self.lib.cf_server_submit(self.h, seq.encode("ascii", "replace"), L,
                                            m.ctypes.data if m is not None else None, k, _c.c_float(gamma),
                                            int(n_samples) if k == CF_SAMPLE else 0, key, 0.0)
            check(int(rid), self.engine.ctx, "cf_server_submit")
        self.meta[int(rid)] = (kind, L, int(n_samples) if k == CF_SAMPLE else 0)
        return int(rid)

    def poll(self, cap=4096):
        ids = (_c.c_int64 * cap)()
        with self.lock:
            n = self.lib.cf_server_poll(self.h, ids, cap)
            check(n, self.engine.ctx, "cf_server_poll")
        return [int(ids[i]) for i in range(n)]

    def pending(self):
        return int(self.lib.cf_server_pending(self.h))

    STATS = ("batches", "folds", "host_launch_ms", "host_retire_ms", "batch_residence_ms", "max_queue")

    def stats(self):
        """Scheduler counters since creation (cf_server_stats)."""
        out = np.zeros(len(self.STATS), np.float64)
        with self.lock:
            n = self.lib.cf_server_stats(self.h, out.ctypes.data_as(_c.POINTER(_c.c_double)), len(out))
            check(min(n, 0), self.engine.ctx, "cf_server_stats")
        d = dict(zip(self.STATS, out.tolist()))
        for k in ("host_launch_ms", "host_retire_ms", "batch_residence_ms"):
            d[k] /= 1e3
        return d

    def result(self, rid):
        kind, L, ns = self.meta.pop(rid)
        size = {"mea": L, "sample": ns * L, "bpp": 4 * L * L, "logz": 0}[kind]
        buf = _c.create_string_buffer(max(size, 1))
        lz = _c.c_float()
        lat = _c.c_double()
        with self.lock:
            check(self.lib.cf_server_result(self.h, rid, buf, _c.byref(lz), _c.byref(lat)), self.engine.ctx,
                  "cf_server_result")
        out = {"logz": float(lz.value), "latency_ms": lat.value / 1e3, "kind": kind}
        raw = buf.raw[:size]
        if kind == "mea":
            out["structure"] = raw.decode("ascii")
        elif kind == "sample":
            s = raw.decode("ascii")
            out["samples"] = [s[t * L:(t + 1) * L] for t in range(ns)]
        elif kind == "bpp":
            out["bpp"] = np.frombuffer(raw, np.float32).reshape(L, L).astype(np.float64)
        return out

    def run_all(self, timeout_s=600.0):
        """Poll until every submitted request is complete; returns {id: result}."""
        res = {}
        t0 = time.time()
        while self.meta:
            for rid in self.poll():
                res[rid] = self.result(rid)
            if time.time() - t0 > timeout_s:
                raise TimeoutError("server did not drain")
        return res

    def replay(self, trace, seed=0, with_wait=False):
        """Replay a trace in real time inside the library's scheduler loop.
        Returns (latency_ms array in trace order, wall seconds), and with
        with_wait the queued part of each latency (ms) as a third item."""
        seqs = trace["seqs"]
        n = len(seqs)
        from .engine import encode_batch
        buf, offs, mbuf, _ = encode_batch(seqs, trace.get("masks"))
        kinds = np.array([KINDS[k] for k in trace["kinds"]], np.uint32)
        gammas = np.full(n, 6.0, np.float32)
        ns = np.array([10 if k == "sample" else 0 for k in trace["kinds"]], np.int32)
        arr = np.ascontiguousarray(np.asarray(trace["arrival_s"], np.float64) * 1e6)
        lat = np.zeros(n, np.float64)
        wait = np.zeros(n, np.float64)
        wall = _c.c_double()
        with self.lock:
            check(self.lib.cf_server_replay(self.h, n, buf, offs.ctypes.data_as(_c.POINTER(_c.c_int64)),
                                            mbuf.ctypes.data if mbuf is not None else None,
                                            kinds.ctypes.data_as(_c.POINTER(_c.c_uint)),
                                            gammas.ctypes.data_as(_c.POINTER(_c.c_float)),
                                            ns.ctypes.data_as(_c.POINTER(_c.c_int)),
                                            arr.ctypes.data_as(_c.POINTER(_c.c_double)), int(seed),
                                            lat.ctypes.data_as(_c.POINTER(_c.c_double)),
                                            wait.ctypes.data_as(_c.POINTER(_c.c_double)), _c.byref(wall)),
                  self.engine.ctx, "cf_server_replay")
        if with_wait:
            return lat / 1e3, wall.value, wait / 1e3
        return lat / 1e3, wall.value


def steady_rate(arrival_s, latency_ms):
    """Completions per second while the middle 80% of the arrivals come in
    (requests sorted by arrival; a negative latency marks a failed request).
    Unlike n / wall, this does not count the drain of the last long folds
    after the final arrival, so a server that keeps up reads ~ the offered rate."""
    arr = np.asarray(arrival_s, np.float64)
    lat = np.asarray(latency_ms, np.float64)
    n = arr.size
    done = arr + np.where(lat >= 0, lat, np.inf) / 1e3
    lo_t, hi_t = arr[n // 10], arr[(9 * n) // 10]
    return float(((done >= lo_t) & (done < hi_t)).sum()) / max(hi_t - lo_t, 1e-9)


def plain_throughput(engine, trace, repeats=2):
    """Folds/s of the same request mix folded as plain batches (one batch per
    kind, inputs resident, steady state)."""
    groups = {}
    for k, (s, m, kind) in enumerate(zip(trace["seqs"], trace["masks"], trace["kinds"])):
        groups.setdefault(kind, []).append((s, m))
    batches = []
    for kind, items in groups.items():
        b = engine.batch([s for s, _ in items], [m for _, m in items])
        batches.append((kind, b))
    def once():
        for kind, b in batches:
            if kind == "mea":
                b.run(CF_MEA, gammas=[6.0], recompute=True)
            elif kind == "sample":
                b.run(CF_SAMPLE, n_samples=10, recompute=True)
            elif kind == "bpp":
                b.run(CF_BPP, recompute=True)
            else:
                b.run(CF_LOGZ, recompute=True)
    once()
    t = time.perf_counter()
    for _ in range(repeats):
        once()
    dt = (time.perf_counter() - t) / repeats
    for _, b in batches:
        b.close()
    return len(trace["seqs"]) / dt


def benchmark(engine, P, n=12000, fracs=(0.25, 0.5, 0.7, 0.85, 1.0), seed=5, log=print, n_streams=16, plain=None,
              seconds=None, n_plain=None):
    """Serving workload: Poisson arrivals at fractions of the plain batched
    capacity of the same mix; p50/p99 latency and served folds/s per rate.

    ``seconds``: if given, each rate's trace spans that many seconds of
    arrivals (n = rate x seconds requests) instead of ``n`` requests, so every
    rate reaches a steady state.  ``n_plain``: requests of the trace whose
    plain batched throughput is the reference (default ``n``)."""
    from . import workloads
    engine.set_params(P)
    base = workloads.serving_trace(n_plain or n, 1.0, seed=seed)
    if plain is None:
        plain = plain_throughput(engine, base)
    log(f"[serve] plain batched throughput of the mix: {plain:.0f} folds/s")
    srv = Server(engine, n_streams=n_streams)
    rows = []
    prev = {}
    try:
        for f in fracs:
            qps = f * plain
            n_f = int(max(500, qps * seconds)) if seconds else n
            tr = workloads.serving_trace(n_f, qps, seed=seed)
            lat, wall, wait = srv.replay(tr, with_wait=True)
            ok = lat >= 0
            sustained = steady_rate(tr["arrival_s"], lat)
            row = {"offered_frac": f, "offered_qps": round(qps, 1), "requests": n_f,
                   "arrival_span_s": round(float(tr["arrival_s"][-1]), 2), "wall_s": round(wall, 2),
                   "served_folds_per_s": round(n_f / wall, 1),
                   "served_vs_plain": round(n_f / wall / plain, 4),
                   "sustained_folds_per_s": round(sustained, 1), "sustained_vs_plain": round(sustained / plain, 4),
                   "p50_ms": round(float(np.percentile(lat[ok], 50)), 3),
                   "p99_ms": round(float(np.percentile(lat[ok], 99)), 3), "failed": int((~ok).sum()),
                   "queued_p50_ms": round(float(np.percentile(wait[ok], 50)), 3)}
            # latency by length class: the long tail's own fold time bounds p99
            L = np.asarray(tr["lengths"])
            by = {}
            run = lat - wait
            for lo, hi in ((0, 100), (100, 200), (200, 400), (400, 700), (700, 1001)):
                m = ok & (L > lo) & (L <= hi)
                if m.any():
                    by[f"{lo + 1}-{hi}"] = {"n": int(m.sum()), "p50_ms": round(float(np.percentile(lat[m], 50)), 1),
                                            "p99_ms": round(float(np.percentile(lat[m], 99)), 1),
                                            "queued_p50_ms": round(float(np.percentile(wait[m], 50)), 1),
                                            "queued_p99_ms": round(float(np.percentile(wait[m], 99)), 1),
                                            "run_p50_ms": round(float(np.percentile(run[m], 50)), 1),
                                            "run_p99_ms": round(float(np.percentile(run[m], 99)), 1)}
            row["by_length"] = by
            st = srv.stats()
            row["server"] = {k: (round(v - prev.get(k, 0.0), 2) if k != "max_queue" else v) for k, v in st.items()}
            prev = st
            rows.append(row)
            log(f"[serve] {row}")
    finally:
        srv.close()
    best = [r for r in rows if r["p99_ms"] <= 50.0]
    head = max(best, key=lambda r: r["sustained_folds_per_s"]) if best else None
    span = f"{seconds:g} s of arrivals per rate" if seconds else f"{n} requests per rate"
    return {"n_streams": n_streams,
            "workload": f"{span}, log-normal lengths (median 150, sigma 0.6, 20-1000 nt), 30% masked "
                        f"(p=0.15), mix 70% MEA / 20% sample x10 / 10% bpp, Poisson arrivals",
            "plain_folds_per_s": round(plain, 1), "rates": rows,
            "best_at_p99_le_50ms": head}
