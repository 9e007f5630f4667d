"""Open-loop load generator for the serving loop (gbsw_server_*).

The serving workload: Poisson arrivals of requests of 1..4096 pairs
(log-uniform), 70% DNA drawn like the 1M-pair DNA mix and 30% protein drawn like
the 20,000-pair protein set; offered load is a fraction of the peak GCUPS the
same server sustains at saturation (closed loop: a fixed number of requests in
flight, each completion replaced at once).

A pool of requests is generated once and laid out in pinned host memory
(gbsw_host_alloc), so submissions are zero-copy and cost the generator only the
offset scan inside gbsw_server_submit.  Arrivals draw pool entries in a seeded
random order; each in-flight submission gets its own result arrays.  Latency is the server's own
submit -> results-written time per request; the generator also reports how far
its submissions fell behind the Poisson schedule (``lag``), which bounds how
faithfully the offered load was applied.
"""
import ctypes
import time

import numpy as np

import gbsw
from gbsw import synth
from gbsw.scoring import AA_GAP_EXTEND, AA_GAP_OPEN, BLOSUM62


class Pool:
    """Requests resident in one pinned host buffer."""

    def __init__(self, n_requests=512, seed=None, max_pairs=4096):
        reqs = synth.serving_requests(n_requests, seed=seed, max_pairs=max_pairs)
        total = sum(len(w.ref) + len(w.query) for w in reqs)
        self.buf = gbsw.HostBuffer(total + 16)
        a = self.buf.array
        self.requests = []
        self.scorings = {gbsw.DNA: gbsw.dna_scoring(*synth.DNA_WORKLOAD_SCORES),
                         gbsw.PROTEIN: gbsw.protein_scoring(BLOSUM62, AA_GAP_OPEN, AA_GAP_EXTEND)}
        at = 0
        for w in reqs:
            nr, nq = len(w.ref), len(w.query)
            a[at:at + nr] = w.ref
            ro = (w.ref_off + np.uint64(at)).astype(np.uint64)
            at += nr
            a[at:at + nq] = w.query
            qo = (w.query_off + np.uint64(at)).astype(np.uint64)
            at += nq
            self.requests.append({"n": w.n, "cells": w.cells(), "alphabet": w.alphabet, "ro": ro, "qo": qo,
                                  "workload": w})
        self.base = a.ctypes.data
        self.max_pairs = max(r["n"] for r in self.requests)
        self.mean_cells = float(np.mean([r["cells"] for r in self.requests]))

    def close(self):
        self.buf.close()


class _Outputs:
    """Free list of 5 x max_pairs int32 result blocks for in-flight requests."""

    def __init__(self, max_pairs):
        self.max_pairs = max(1, max_pairs)
        self.free = []

    def get(self):
        return self.free.pop() if self.free else np.zeros((5, self.max_pairs), dtype=np.int32)

    def put(self, o):
        self.free.append(o)


def _submit_fn(server, pool, outputs):
    lib = server._lib
    h = server._h
    base = pool.base
    scp = {k: ctypes.byref(v) for k, v in pool.scorings.items()}
    ticket = ctypes.c_uint64()
    tref = ctypes.byref(ticket)

    def submit(req):
        out = outputs.get()
        p = out.ctypes.data
        st = out.strides[0]
        rc = lib.gbsw_server_submit(h, scp[req["alphabet"]], req["n"], base, req["ro"].ctypes.data, base,
                                    req["qo"].ctypes.data, p, p + st, p + 2 * st, p + 3 * st, p + 4 * st, tref)
        if rc != 0:
            raise gbsw.GbswError(rc, lib.gbsw_last_error().decode(errors="replace"))
        return ticket.value, out

    return submit


def saturate(server, pool, repeat=1):
    """Submit the whole pool at once (repeat times); returns the sustained GCUPS
    from first submit to last completion, and the per-request infos."""
    outputs = _Outputs(pool.max_pairs)
    submit = _submit_fn(server, pool, outputs)
    inflight = {}
    infos = []
    cells = 0
    t0 = time.perf_counter()
    for _ in range(repeat):
        for req in pool.requests:
            t, out = submit(req)
            inflight[t] = out
            cells += req["cells"]
    while inflight:
        done = server.poll()
        if not done:
            time.sleep(0.0002)
        for info in done:
            outputs.put(inflight.pop(info["ticket"]))
            infos.append(info)
    dt = time.perf_counter() - t0
    return cells / dt / 1e9, infos


def closed_loop(server, pool, duration, depth=None, warmup=0.5, seed=0):
    """Sustained throughput at saturation: `depth` requests (default: two pools'
    worth) stay in flight, each completion is replaced by the next pool entry
    (seeded random order), and the GCUPS is the cells of the requests that
    complete within `duration` seconds after a `warmup`, over that window.
    Returns (gcups, completed requests in the window)."""
    rng = np.random.default_rng(seed)
    depth = depth or 2 * len(pool.requests)
    outputs = _Outputs(pool.max_pairs)
    submit = _submit_fn(server, pool, outputs)
    order = rng.permutation(len(pool.requests))
    nxt = 0
    inflight = {}

    def refill():
        nonlocal nxt
        while len(inflight) < depth:
            req = pool.requests[order[nxt % len(order)]]
            t, out = submit(req)
            inflight[t] = (out, req["cells"])
            nxt += 1

    refill()
    t0 = time.perf_counter()
    t_on, t_off = t0 + warmup, t0 + warmup + duration
    cells = 0
    count = 0
    while True:
        now = time.perf_counter()
        if now >= t_off:
            break
        done = server.poll()
        if not done:
            time.sleep(0.0002)
            continue
        for info in done:
            out, c = inflight.pop(info["ticket"])
            outputs.put(out)
            if t_on <= now < t_off:
                cells += c
                count += 1
        refill()
    while inflight:  # drain
        for info in server.poll():
            outputs.put(inflight.pop(info["ticket"])[0])
        time.sleep(0.0002)
    return cells / duration / 1e9, count


def poisson(server, pool, rate, duration, seed=0, verify=None):
    """Offer `rate` requests/s for `duration` seconds (Poisson arrivals, pool
    entries in a seeded random order).  verify(pool_index, out, info) is called
    on completed requests when given.  Returns a summary dict."""
    rng = np.random.default_rng(seed)
    n_arr = max(1, int(rate * duration * 1.5) + 16)
    gaps = rng.exponential(1.0 / rate, size=n_arr)
    arrivals = np.cumsum(gaps)
    arrivals = arrivals[arrivals < duration]
    order = rng.integers(0, len(pool.requests), size=len(arrivals))
    outputs = _Outputs(pool.max_pairs)
    submit = _submit_fn(server, pool, outputs)
    inflight = {}
    infos = []
    lags = np.zeros(len(arrivals))
    cells = 0
    t0 = time.perf_counter()
    i = 0
    last_poll = t0
    while i < len(arrivals):
        now = time.perf_counter() - t0
        if now >= arrivals[i]:
            req = pool.requests[order[i]]
            t, out = submit(req)
            inflight[t] = (out, order[i])
            lags[i] = now - arrivals[i]
            cells += req["cells"]
            i += 1
            continue
        if now - (last_poll - t0) > 0.0005:
            for info in server.poll():
                out, k = inflight.pop(info["ticket"])
                if verify is not None:
                    verify(k, out, info)
                outputs.put(out)
                infos.append(info)
            last_poll = time.perf_counter()
    t_sub = time.perf_counter() - t0
    while inflight:
        done = server.poll()
        if not done:
            time.sleep(0.0002)
        for info in done:
            out, k = inflight.pop(info["ticket"])
            if verify is not None:
                verify(k, out, info)
            outputs.put(out)
            infos.append(info)
    t_end = time.perf_counter() - t0
    lat = np.array([x["latency_ms"] for x in infos])
    bad = sum(1 for x in infos if x["status"] != 0)
    return {
        "offered_req_per_s": rate,
        "requests": len(infos),
        "errors": bad,
        "cells": cells,
        "window_s": duration,
        "gcups": cells / max(t_end, 1e-9) / 1e9,
        "offered_gcups": rate * pool.mean_cells / 1e9,
        "p50_ms": float(np.percentile(lat, 50)) if len(lat) else None,
        "p99_ms": float(np.percentile(lat, 99)) if len(lat) else None,
        "max_ms": float(lat.max()) if len(lat) else None,
        "lag_p99_ms": float(np.percentile(lags, 99) * 1e3) if len(lags) else None,
        "submit_s": t_sub,
        "drain_s": t_end - t_sub,
    }
