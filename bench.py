#!/usr/bin/env python3
"""Benchmark of the batched Smith-Waterman hot path on one MI355X.

A "step" is one full alignment (forward + reverse, all five outputs) of the
1,000,000-pair uneven-length DNA mix (config 2) already resident in HBM, through
the C ABI (gbsw_batch_align).  The headline value is GCUPS of the batched form:
sum(|ref|*|query|) per step / wall seconds (reference convention: forward cells
counted, forward + reverse timed).  The same batch is also timed in the plain
form (one forward and one reverse launch, input order, one geometry), and the
CPU oracle is timed on a bounded sample of the same workload.

Serving (the "p99 latency under load" half of the metric): a server
(gbsw_server_*) is driven by an open-loop Poisson generator over a pool of
requests of 1..4096 pairs (70% DNA like the mix above, 30% protein); its peak is
measured by submitting the whole pool at once, then each offered load (25, 50,
75, 100% of that peak) runs for --serving-window seconds and reports p50/p99
request latency and sustained GCUPS.  Every completed request is compared with
the same request aligned alone.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "gpu-bsw_amd"))

# Hardware queues per process, read once when HIP initialises (before any GPU
# call here).  The serving loop runs four slots of four streams; with HIP's
# default of four queues their kernels serialise behind each other's long
# buckets (measured: p50 latency at 25 % load 26 ms with 4 queues, 7 ms with 16).
os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")

import numpy as np  # noqa: E402

import gbsw  # noqa: E402
from gbsw import synth  # noqa: E402

# MI355X: 256 CUs, 2.4 GHz peak engine clock, 64 lanes per CU per clock for
# 32-bit integer VALU; packed int16 doubles the cell rate.  12 integer ops per
# cell update (E: add+max, F: add+max, shared H+open, substitution, diagonal add,
# max3+max0, best tracking cmp+max+select).
MI355X_CU, MI355X_GHZ, OPS_PER_CELL = 256, 2.4, 12
DEFAULT_GEOMETRY = 1  # the library's default length -> geometry map (GBSW_KMODEL unset)
PEAK_INT_TOPS_PACKED = MI355X_CU * 64 * MI355X_GHZ * 2 / 1000.0  # 78.6 TOP/s (int16x2)


# launch kind -> (kernel, cell policy)
KIND_NAMES = {2: ("k_batched", "Packed16"), 3: ("k_batched", "Wide32"), 4: ("k_batched", "PackedF16"),
              5: ("k_split", "Packed16"), 6: ("k_split", "PackedF16")}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profiled_traffic(kname):
    """HBM bytes per launch of `kname` (FETCH_SIZE + WRITE_SIZE, calibrated) from
    the committed rocprofv3 summary, or None when it does not cover this kernel."""
    import re
    path = os.path.join(REPO, "profiles", "k_batched_dna_mix_1m.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"\s*traffic_bytes_per_launch\s+(.+?)\s*=\s*([0-9.eE+]+)", line)
        if m and m.group(1).strip() == kname:
            return float(m.group(2))
    return None


def cpu_baseline(w, sample_pairs, threads, device_res):
    """The oracle on the first sample_pairs pairs of the workload (timed), and
    its results checked against the device's for the same pairs."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_lib as ol
    sub = w.subset(np.arange(min(sample_pairs, w.n)))
    sc = ol.dna_scoring(*w.scoring)
    nt = ol.load().gbsw_oracle_threads(threads)
    t0 = time.perf_counter()
    exp = ol.align_batch(sub.ref, sub.ref_off, sub.query, sub.query_off, sc, nthreads=nt)
    dt = time.perf_counter() - t0
    differ = np.zeros(sub.n, dtype=bool)
    for f in gbsw.FIELDS:
        differ |= device_res[f][:sub.n] != exp[f]
    return {"value": round(sub.cells() / dt / 1e9, 4), "unit": "GCUPS", "cores": int(nt), "kind": "port",
            "sample": "first %d pairs of the workload (%.3g cells), oracle fwd+rev, %.1f s" % (sub.n, sub.cells(), dt),
            "bit_exact_vs_oracle_pairs": int(sub.n - differ.sum()), "mismatched_pairs": int(differ.sum())}


def traceback_bench(batch, scoring, res, w, args):
    """Alignment paths (gbsw_batch_traceback) of the aligned batch: device and wall
    time of the whole batch, in rectangle cells per second (the rectangle of a
    pair: ref_end - ref_begin by query_end - query_begin)."""
    on = res["score"] > 0
    rows = (res["ref_end"] - res["ref_begin"]).astype(np.int64)
    cols = (res["query_end"] - res["query_begin"]).astype(np.int64)
    rect = int(np.sum(np.where(on, rows * cols, 0)))
    cap = int(np.sum(np.where(on, rows + cols, 0)))
    batch.traceback(scoring, capacity=cap)  # warm-up (scratch allocation)
    dev, wall = [], []
    for _ in range(max(1, args.traceback_steps)):
        t = time.perf_counter()
        words, off, ps, ms = batch.traceback(scoring, capacity=cap)
        wall.append(time.perf_counter() - t)
        dev.append(ms)
    d, wl = min(dev), min(wall)
    out = {"pairs": int(w.n), "rect_cells": rect, "device_ms": round(d, 3), "wall_ms": round(1e3 * wl, 3),
           "device_gcups_rect": round(rect / d / 1e6, 2), "wall_gcups_rect": round(rect / wl / 1e9, 2),
           "cigar_words": int(off[-1]), "path_score_equals_score": round(float(np.mean(ps[on] == res["score"][on])), 6)}
    log("traceback: %d pairs, %.4g rectangle cells, device %.2f ms (%.1f GCUPS), wall %.2f ms, %d CIGAR words" %
        (w.n, rect, d, out["device_gcups_rect"], 1e3 * wl, out["cigar_words"]))
    if not args.no_cpu:
        # CPU baseline: the oracle's paths on a bounded sample, checked against the
        # device's words for the same pairs
        sys.path.insert(0, os.path.join(REPO, "tests"))
        import oracle_lib as ol
        k = min(args.traceback_cpu_pairs, w.n)
        sub = w.subset(np.arange(k))
        r = np.zeros(k, dtype=ol.RESULT_DTYPE)
        for f in gbsw.FIELDS:
            r[f] = res[f][:k]
        nt = ol.load().gbsw_oracle_threads(args.cpu_threads)
        t = time.perf_counter()
        ew, eoff, eps = ol.traceback_batch(sub.ref, sub.ref_off, sub.query, sub.query_off, ol.dna_scoring(*w.scoring),
                                           r, nthreads=nt)
        dt = time.perf_counter() - t
        srect = int(np.sum(np.where(on[:k], rows[:k] * cols[:k], 0)))
        same = np.array_equal(ew, words[:off[k]]) and np.array_equal(eoff, off[:k + 1]) and np.array_equal(eps, ps[:k])
        out["cpu_baseline"] = {"value": round(srect / dt / 1e9, 4), "unit": "GCUPS (rectangle cells)", "cores": int(nt),
                               "kind": "port", "sample": "first %d pairs (%.3g rectangle cells), oracle paths, %.1f s" %
                               (k, srect, dt)}
        out["bit_exact_sample_vs_oracle"] = bool(same)
        log("traceback cpu oracle: %.3f GCUPS on %d threads, first %d pairs bit-exact: %s" %
            (srect / dt / 1e9, nt, k, same))
    return out


def dropin_bench(w, res, args):
    """End-to-end calls of the C++ drop-in (gpu_bsw_driver::kernel_driver_dna) on the
    workload, timed the way the reference reports it: its driver's "Total Execution
    Time" line (packing, transfers, both passes, copy back; driver.cpp:34, 240-242).
    The sums of scores and coordinates of the last call must equal the batch's."""
    import shutil
    import subprocess
    import tempfile
    exe = os.path.join(REPO, "gpu-bsw_amd", "dropin_time")
    if not os.path.exists(exe):
        subprocess.run(["make", "-s", "-C", os.path.join(REPO, "gpu-bsw_amd"), "dropin_time"], check=True)
    tmp = tempfile.mkdtemp(prefix="gbsw_dropin_")
    try:
        fr, fq = os.path.join(tmp, "refs.txt"), os.path.join(tmp, "queries.txt")
        for path, buf, off in ((fr, w.ref, w.ref_off), (fq, w.query, w.query_off)):
            # one sequence per line: the bytes with a newline after each pair's end
            lines = np.insert(np.asarray(buf, dtype=np.uint8), np.asarray(off[1:], dtype=np.int64), ord("\n"))
            lines.tofile(path)
        out = subprocess.run([exe, fr, fq, str(args.dropin_reps + 1)], capture_output=True, text=True, check=True,
                             timeout=600).stdout.splitlines()
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    times = [float(l.split(":")[1]) for l in out if l.startswith("Total Execution Time (seconds):")]
    ck = [l.split() for l in out if l.startswith("checksum")][0]
    want_score = int(np.sum(res["score"], dtype=np.int64))
    want_coord = int(np.sum(res["ref_begin"], dtype=np.int64) + np.sum(res["ref_end"], dtype=np.int64) +
                     np.sum(res["query_begin"], dtype=np.int64) + np.sum(res["query_end"], dtype=np.int64))
    same = int(ck[1]) == want_score and int(ck[2]) == want_coord
    t = min(times[1:]) if len(times) > 1 else times[0]
    log("drop-in end to end: %d pairs, Total Execution Time %.4f s (best of %d after one warm-up call), "
        "%.1f GCUPS; sums equal to the batch's: %s" % (w.n, t, len(times) - 1, w.cells() / t / 1e9, same))
    return {"pairs": int(w.n), "total_execution_time_s": round(t, 4), "gcups": round(w.cells() / t / 1e9, 2),
            "calls": len(times), "sums_equal_batch": bool(same)}


def time_mode(batch, scoring, mode, warmup, steps, cells):
    for _ in range(warmup):
        batch.align(scoring, mode)
    device_ms, launches = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.align(scoring, mode)
        st = batch.stats()
        device_ms.append(st["device_ms"])
        launches.append(batch.launches())
    wall = time.perf_counter() - t0
    # per-launch durations averaged over the timed steps
    avg = []
    for i, L in enumerate(launches[0]):
        d = dict(L)
        d["ms"] = float(np.mean([step[i]["ms"] for step in launches]))
        avg.append(d)
    return {"wall_s": wall, "steps": steps, "ms_per_step": 1000.0 * wall / steps, "gcups": cells * steps / wall / 1e9,
            "device_ms": float(np.mean(device_ms)), "device_gcups": cells / (np.mean(device_ms) / 1000.0) / 1e9,
            "launches": avg, "buckets": st["buckets"]}


def steps_for(batch, scoring, mode, floor, target_s=0.3, cap=500):
    """Timed steps for a config: enough for about target_s of work (a 10k-pair
    call takes half a millisecond, so a few steps would time host jitter),
    at least floor; the align it times is one more warm-up."""
    t = time.perf_counter()
    batch.align(scoring, mode)
    dt = time.perf_counter() - t
    return max(floor, min(cap, int(target_s / max(dt, 1e-6))))


def config_bench(args, only=None):
    """The other workloads: 10k plain DNA and 20k protein (plain and batched
    forms, bit-exact between them) and the 50k long-protein set (batched); with
    `only`, that one workload's batched form (a profiling run)."""
    from gbsw.scoring import BLOSUM62
    out = {}
    ctx = gbsw.Context(0)
    for name, make, plain in (("plain_dna_10k", lambda: synth.plain_dna(), True),
                              ("plain_protein_20k", lambda: synth.plain_protein(), True),
                              ("long_protein", lambda: synth.long_protein(args.long_pairs), False)):
        if only is not None:
            if name != only:
                continue
            plain = False
        t = time.perf_counter()
        w = make()
        cells = w.cells()
        gen_s = time.perf_counter() - t
        sc = gbsw.dna_scoring(*w.scoring) if w.alphabet == gbsw.DNA else gbsw.protein_scoring(BLOSUM62, *w.scoring)
        b = ctx.batch(w.alphabet, w.ref, w.ref_off, w.query, w.query_off)
        # a profiling run (`only`) keeps one warm-up and two timed aligns
        steps = 2 if only is not None else steps_for(b, sc, "batched", 3 if plain else 2)
        rb = time_mode(b, sc, "batched", 1, steps, cells)
        res_b = b.results()
        d = {"pairs": w.n, "cells": cells, "batched_steps": rb["steps"], "batched_gcups": round(rb["gcups"], 2),
             "batched_device_gcups": round(rb["device_gcups"], 2), "batched_ms": round(rb["ms_per_step"], 3),
             "int32_redone_pairs": int(sum(L["pairs"] for L in rb["launches"]
                                           if L["kind"] & 0xFF == gbsw.KIND_BATCHED_WIDE)),
             # the whole step against the packed int16 VALU roofline (same op model as the
             # headline's aggregate fraction)
             "roofline_frac": round(OPS_PER_CELL * cells / (rb["ms_per_step"] * 1e-3) / 1e12 / PEAK_INT_TOPS_PACKED, 4)}
        if plain:
            rp = time_mode(b, sc, "plain", 1, steps_for(b, sc, "plain", 2), cells)
            res_p = b.results()
            d.update({"plain_steps": rp["steps"], "plain_gcups": round(rp["gcups"], 2), "plain_ms": round(rp["ms_per_step"], 3),
                      "batched_over_plain": round(rb["gcups"] / rp["gcups"], 3),
                      "bit_exact_plain_vs_batched": all(np.array_equal(res_b[f], res_p[f]) for f in gbsw.FIELDS)})
        log("%s: %d pairs, %.4g cells (generated in %.1f s): batched %.1f GCUPS%s" %
            (name, w.n, cells, gen_s, rb["gcups"],
             ", plain %.1f GCUPS, bit-exact %s" % (d["plain_gcups"], d["bit_exact_plain_vs_batched"]) if plain else ""))
        out[name] = d
        b.close()
        del w
    ctx.close()
    return out


def serving_bench(args):
    from gbsw import loadgen
    t = time.perf_counter()
    pool = loadgen.Pool(args.serving_requests)
    log("serving pool: %d requests, %d pairs, mean %.3g cells/request, built in %.1f s" %
        (len(pool.requests), sum(r["n"] for r in pool.requests), pool.mean_cells, time.perf_counter() - t))
    # each request aligned alone: the results every served copy must reproduce.
    # The context is closed before the server starts, so that the server's
    # streams have the hardware queues to themselves.
    alone = []
    ctx = gbsw.Context(0)
    for r in pool.requests:
        w = r["workload"]
        res = ctx.align_packed(w.ref, w.ref_off, w.query, w.query_off, pool.scorings[w.alphabet], "batched")
        alone.append(np.stack([res[f] for f in gbsw.FIELDS]))
    ctx.close()
    mismatches = [0]

    def verify(k, out, info):
        n = alone[k].shape[1]
        if info["status"] != 0 or not np.array_equal(out[:, :n], alone[k]):
            mismatches[0] += 1

    srv = gbsw.Server(0, slots=args.serving_slots, pairs_per_slot=args.serving_pairs_per_slot)
    loadgen.saturate(srv, pool, repeat=1)  # warm-up
    burst, _ = loadgen.saturate(srv, pool, repeat=2)
    # peak: saturation in closed loop (four pools' worth of requests in flight,
    # each completion replaced at once), the rate the offered loads are set from
    depth = 4 * len(pool.requests)
    peak, peak_requests = loadgen.closed_loop(srv, pool, args.serving_peak_window, depth=depth)
    log("serving peak (closed loop, %d requests in flight, %.0f s): %.1f GCUPS (%d requests); whole pool x2 "
        "at once: %.1f GCUPS" % (depth, args.serving_peak_window, peak, peak_requests, burst))
    loads = []
    for frac in [float(x) for x in args.serving_loads.split(",")]:
        rate = frac * peak * 1e9 / pool.mean_cells
        r = loadgen.poisson(srv, pool, rate, args.serving_window, seed=int(frac * 1000), verify=verify)
        log("  load %.0f%%: offered %.1f GCUPS (%.0f req/s), sustained %.1f GCUPS, p50 %.2f ms, p99 %.2f ms, "
            "%d requests, submit lag p99 %.2f ms" % (100 * frac, r["offered_gcups"], rate, r["gcups"], r["p50_ms"],
                                                     r["p99_ms"], r["requests"], r["lag_p99_ms"]))
        loads.append({"load": frac, "offered_gcups": round(r["offered_gcups"], 1), "gcups": round(r["gcups"], 1),
                      "req_per_s": round(rate, 1), "requests": r["requests"], "p50_ms": round(r["p50_ms"], 3),
                      "p99_ms": round(r["p99_ms"], 3), "submit_lag_p99_ms": round(r["lag_p99_ms"], 3),
                      "errors": r["errors"]})
    st = srv.stats()
    srv.close()
    pool.close()
    return {"peak_gcups": round(peak, 1), "peak_method": "closed loop, %d requests in flight, %.0f s" % (
                depth, args.serving_peak_window),
            "burst_gcups": round(burst, 1), "window_s": args.serving_window, "slots": args.serving_slots,
            "pairs_per_slot": args.serving_pairs_per_slot,
            "pool_requests": len(pool.requests), "loads": loads, "micro_batches": st["micro_batches"],
            "requests": st["completed"], "hbm_bytes": st["device_bytes"],
            "bit_exact_vs_alone": mismatches[0] == 0, "mismatches": mismatches[0]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pairs", type=int, default=1_000_000, help="pairs of the DNA mix (default: full config)")
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--cpu-pairs", type=int, default=200_000)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="skip the 10k DNA / 20k protein / long-protein lines")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default length -> geometry map")
    ap.add_argument("--geometry", type=int, default=-1,
                    help="use this length -> geometry map (0-5) instead of autotuning one")
    ap.add_argument("--long-pairs", type=int, default=50_000, help="pairs of the long-protein set")
    ap.add_argument("--no-traceback", action="store_true", help="skip the alignment-path (CIGAR) line")
    ap.add_argument("--no-dropin", action="store_true", help="skip the end-to-end drop-in calls")
    ap.add_argument("--dropin-reps", type=int, default=3)
    ap.add_argument("--traceback-steps", type=int, default=3)
    ap.add_argument("--traceback-cpu-pairs", type=int, default=50_000)
    ap.add_argument("--serving-window", type=float, default=60.0, help="seconds per offered load")
    ap.add_argument("--serving-loads", default="0.25,0.5,0.75,1.0")
    ap.add_argument("--serving-requests", type=int, default=512, help="request pool size")
    ap.add_argument("--serving-slots", type=int, default=2)
    ap.add_argument("--serving-pairs-per-slot", type=int, default=524288)
    ap.add_argument("--serving-peak-window", type=float, default=3.0, help="seconds of the closed-loop peak")
    ap.add_argument("--only", choices=["plain_dna_10k", "plain_protein_20k", "long_protein"],
                    help="time only this workload's batched form and print its figures (a profiling run)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run one device-to-device copy of this size (rocprof FETCH/WRITE calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path here (pairs are independent; "
              "several GPUs would each take a share of the pairs balanced by cell count). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    if args.only:
        print(json.dumps({"only": args.only, **config_bench(args, only=args.only)[args.only]}))
        return
    t = time.perf_counter()
    w = synth.batched_dna(args.pairs)
    cells = w.cells()
    log("workload %s: %d pairs, %.4g cells, generated in %.1f s" % (w.name, w.n, cells, time.perf_counter() - t))

    if args.calib_copy_mib:
        import torch
        src = torch.empty(args.calib_copy_mib << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copy: %d MiB device-to-device" % args.calib_copy_mib)
        del src, dst

    ctx = gbsw.Context(0)
    tuned = None
    if args.geometry >= 0:
        # (a batch takes its context's map when it is created)
        ctx.set_option(gbsw.OPT_GEOMETRY, args.geometry)
        tuned = {"geometry_model": args.geometry, "pinned": True}
        log("geometry model %d (pinned)" % args.geometry)
    batch = ctx.batch(w.alphabet, w.ref, w.ref_off, w.query, w.query_off)
    scoring = gbsw.dna_scoring(*w.scoring)
    if args.geometry < 0 and not args.no_autotune:
        # untimed set-up: the length -> geometry map measured on this batch
        t = time.perf_counter()
        model, model_ms = batch.autotune(scoring, repeats=3)
        tuned = {"geometry_model": model, "device_ms": round(model_ms, 3), "tuning_s": round(time.perf_counter() - t, 2)}
        log("autotune: geometry model %d (%.2f ms device), %.1f s" % (model, model_ms, tuned["tuning_s"]))

    batched = time_mode(batch, scoring, "batched", args.warmup, args.steps, cells)
    log("batched: %.1f GCUPS wall, %.1f GCUPS device, %.2f ms/step, %d launches" %
        (batched["gcups"], batched["device_gcups"], batched["ms_per_step"], len(batched["launches"])))
    for L in sorted(batched["launches"], key=lambda x: -x["ms"]):
        if L["ms"] > 0:  # (graph replays do not time single launches)
            log("   K=%2d  %7d pairs  %.3g cells  %7.3f ms  %7.1f GCUPS  grid %d" %
                (L["columns_per_lane"], L["pairs"], L["cells"], L["ms"], L["cells"] / L["ms"] / 1e6, L["workgroups"]))
    res_b = batch.results()
    plain = None
    same_plain = None
    if not args.no_plain:
        plain = time_mode(batch, scoring, "plain", 1, args.plain_steps, cells)
        log("plain:   %.1f GCUPS wall, %.1f GCUPS device, %.2f ms/step" %
            (plain["gcups"], plain["device_gcups"], plain["ms_per_step"]))
        res_p = batch.results()
        same = same_plain = all(np.array_equal(res_b[f], res_p[f]) for f in gbsw.FIELDS)
        if not same:
            log("ERROR: plain and batched results differ")
            sys.exit(1)

    # Dominant kernel.  The timed steps run the length buckets on four streams, so
    # their launches share the GPU and a launch's duration is not its own.  One
    # more step runs the buckets one after another and times each launch (one
    # length bucket, persistent, forward and reverse fused); a launch's
    # algorithmic work = 12 integer ops x the bucket's forward cells (|ref| x |query|).
    # (issued call by call: per-launch durations come from hipEvents around
    # each launch, which a graph replay does not time)
    ctx.set_option(gbsw.OPT_STREAMS, 1)
    ctx.set_option(gbsw.OPT_GRAPHS, 0)
    serial = time_mode(batch, scoring, "batched", 1, 1, cells)
    ctx.set_option(gbsw.OPT_STREAMS, 4)
    ctx.set_option(gbsw.OPT_GRAPHS, 1)
    # The dominant kernel is the kernel (template instance) that aligns the most
    # forward cells of the step (fixed by the workload and the length -> geometry
    # map; the launch durations of the two largest buckets lie within a few per
    # cent of each other, so choosing by time would flip between them from run to
    # run).  One kernel may serve several buckets (a single-strip and a
    # multi-strip one), so its duration is the average over its launches, as
    # rocprofv3's kernel statistics give it.
    per_kernel = {}
    for L in serial["launches"]:
        kfn, kpol = KIND_NAMES.get(L["kind"] & 0xFF, ("?", "?"))
        name = "%s<%s, %d, %s>" % (kfn, kpol, L["columns_per_lane"],
                                   "true" if L["kind"] & gbsw.KIND_BYTE_PROFILE else "false")
        d = per_kernel.setdefault(name, {"ms": 0.0, "cells": 0, "pairs": 0, "launches": 0})
        d["ms"] += L["ms"]
        d["cells"] += L["cells"]
        d["pairs"] += L["pairs"]
        d["launches"] += 1
    kname, dom = max(per_kernel.items(), key=lambda kv: kv[1]["cells"])
    dom_ms = dom["ms"] / dom["launches"]
    dom_cells = dom["cells"] / dom["launches"]
    achieved = OPS_PER_CELL * dom_cells / (dom_ms / 1000.0) / 1e12
    agg = OPS_PER_CELL * cells / (batched["device_ms"] / 1000.0) / 1e12
    log("dominant kernel (buckets serialised, %.2f ms/step): %s, %d launch(es), %d pairs, %.4g cells per launch, "
        "%.3f ms per launch, %.1f GCUPS" % (serial["device_ms"], kname, dom["launches"], dom["pairs"], dom_cells,
                                           dom_ms, dom_cells / dom_ms / 1e6))
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": round(PEAK_INT_TOPS_PACKED, 1),
                "unit": "TOP/s", "frac": round(achieved / PEAK_INT_TOPS_PACKED, 4),
                "traffic": profiled_traffic(kname),
                "kernel": "%s (%d launch(es) per step, %d pairs, %.4g cells per launch)" % (
                    kname, dom["launches"], dom["pairs"], dom_cells),
                "kernel_ms": round(dom_ms, 3), "ops_per_cell": OPS_PER_CELL,
                "geometry_model": tuned["geometry_model"] if tuned else DEFAULT_GEOMETRY,
                "aggregate_achieved": round(agg, 3), "aggregate_frac": round(agg / PEAK_INT_TOPS_PACKED, 4)}

    tb = None
    if not args.no_traceback:
        tb = traceback_bench(batch, scoring, res_b, w, args)
        if tb.get("bit_exact_sample_vs_oracle") is False:
            log("ERROR: device alignment paths differ from the oracle's")
            sys.exit(1)

    batch.close()
    ctx.close()
    dropin = None
    if not args.no_dropin:
        dropin = dropin_bench(w, res_b, args)
        if not dropin["sums_equal_batch"]:
            log("ERROR: the drop-in's results differ from the batch's")
            sys.exit(1)
    configs = None
    if not args.no_configs:
        configs = config_bench(args)
        bad = [k for k, v in configs.items() if v.get("bit_exact_plain_vs_batched") is False]
        if bad:
            log("ERROR: plain and batched results differ on %s" % ", ".join(bad))
            sys.exit(1)
    serving = None
    if not args.no_serving:
        serving = serving_bench(args)
        if not serving["bit_exact_vs_alone"]:
            log("ERROR: %d served requests differ from the same requests aligned alone" % serving["mismatches"])
            sys.exit(1)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(w, args.cpu_pairs, args.cpu_threads, res_b)
        log("cpu oracle: %s GCUPS on %d threads (%s); device results equal the oracle's on %d of its pairs" %
            (cpu["value"], cpu["cores"], cpu["sample"], cpu["bit_exact_vs_oracle_pairs"]))
        if cpu["mismatched_pairs"]:
            log("ERROR: %d pairs of the oracle sample differ from the device results" % cpu["mismatched_pairs"])
            sys.exit(1)

    line = {
        "metric": "GCUPS fwd+rev at 1 GPU, batched vs plain, bit-exact; p99 latency under load",
        "value": round(batched["gcups"], 2),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(batched["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        # int16 cells; length buckets whose score bound is below 2048 compute in
        # packed f16 holding integers (exact: every value is an integer below 2048)
        "dtype": "int16 (f16 integer-exact buckets)",
        "data": "synthetic",
        "config": {"workload": "batched uneven-length DNA mix: %d seeded pairs, query lognormal(median 150) "
                               "[32,1024], ref lognormal(median 600) [64,2048], +3/-3/-6/-1" % w.n,
                   "pairs": w.n, "cells": cells, "form": "batched", "autotune": tuned},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "device_gcups": round(batched["device_gcups"], 2),
        "plain": None if plain is None else {"gcups": round(plain["gcups"], 2),
                                              "device_gcups": round(plain["device_gcups"], 2),
                                              "ms_per_step": round(plain["ms_per_step"], 3)},
        "batched_over_plain": None if plain is None else round(batched["gcups"] / plain["gcups"], 3),
        "bit_exact_plain_vs_batched": same_plain,
        "configs": configs,
        "serving": serving,
        "traceback": tb,
        "dropin_end_to_end": dropin,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
