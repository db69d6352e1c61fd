#!/usr/bin/env python3
"""Per-slide latency of the MI355X RTOD path on the 1-D random-walk stream
W=100k, S=5k, R=0.45, K=50 (1,048,572 points), one JSON line on stdout.

A step is one slide: new S points in, the window's exact outlier list out in
host memory.  `value` is the mean per-slide wall time in microseconds that the
library measures from rtod_slide_device entry (the stream already resident
in HBM) to the sorted list in host memory (lower is better), averaged over
--steps slides after --warmup slides.  The same run through rtod_slide, with
each slide copied from host memory inside the call, is reported beside it as
pcie_inclusive_us_per_slide.

Extra objects on the line:
  roofline      dominant kernel of the slide (HIP events on the library's
                stream, separate pass with per-phase events) against the
                8 TB/s HBM peak, algorithmic bytes B = 4*d*(W+S) + 4*|O|
  cpu_baseline  the grid oracle (same algorithm, C, pthreads) timed on this
                host's cores on a bounded sample of the same stream
  batched       256 independent streams of uneven dim/W/S/R/K (the batched
                workload): per-step time of the batched API after its
                autotuner, its roofline over all streams' bytes, and the grid
                oracle over the same streams on the host cores
  learned       the same workloads with the learned work scheduler (online-
                trained MLP ordering the exact range counts): per-slide time,
                speed-up over the plain schedule, predictor miss rate, slides
                that fell back to the plain schedule, and the number of slides
                whose outlier list differs from the plain schedule's (must be 0)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "μs per slide (W=100k,S=5k,K=50) plain vs learned-scheduled; exact outliers"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def _workloads():
    from rtod_amd import synth

    # name: (dim, N, W, S, R, K, generator, seed)
    return {w.name: (w.dim, w.N, w.W, w.S, w.R, w.K, w.kind, w.seed)
            for w in synth.WORKLOADS.values()}


WORKLOADS = _workloads()


def stream(name: str, slides: int) -> np.ndarray:
    from rtod_amd import synth

    # the measured streams have N points; longer runs extend the same seeded
    # generator
    dim, N, W, S = WORKLOADS[name][:4]
    return synth.WORKLOADS[name].data(W + slides * S)


def run_gpu(xs, W, S, R, K, warmup, steps, flags=0, sched=0, device_input=True, index=0,
            keep_lists=False):
    """device_input: the stream sits in HBM (uploaded once, untimed) and each
    slide is fed by device pointer; else each slide's points come from host
    memory through the pinned staging buffer (PCIe inside the timed call)."""
    import rtod_amd

    dim = xs.shape[1]
    det = rtod_amd.Detector(dim, W, S, R, K, flags=flags, sched_kind=sched, index_kind=index)
    ds = rtod_amd.DeviceStream(xs[: W + (warmup + steps) * S]) if device_input else None
    det.prime(xs[:W])
    totals, outs, lists = [], [], []
    t_outer = 0.0
    st0 = None
    for t in range(1, warmup + steps + 1):
        row = W + (t - 1) * S
        t0 = time.perf_counter()
        o = det.slide_device(ds.at(row)) if ds else det.slide(xs[row: row + S])
        dt = time.perf_counter() - t0
        if t == warmup:
            st0 = det.stats()
        if t > warmup:
            totals.append(det.times.total)
            outs.append(len(o))
            t_outer += dt
            if keep_lists:
                lists.append(o.copy())
    res = {"lib_ms": float(np.mean(totals)), "outer_ms": t_outer / max(1, steps) * 1e3,
           "outliers_mean": float(np.mean(outs)) if outs else 0.0,
           "kernels": det.kernel_times(), "stats": det.stats(), "stats_warm": st0,
           "lists": lists}
    det.close()
    if ds:
        ds.close()
    return res


def run_gpu_pipelined(xs, W, S, R, K, warmup, steps, sched=0, index=0, flags=0):
    """The same slides through rtod_slides_device: slide t+1's new batch is
    sorted by extra workgroups of slide t's launches, up to four slides are
    queued and the host collects each list as it completes.  Stream in HBM;
    `warmup` slides in one call, then `steps` slides in one timed call.
    flags=FLAG_PHASE_TIMES: HIP events around each kernel (a separate,
    slide-by-slide measuring pass)."""
    import rtod_amd

    dim = xs.shape[1]
    det = rtod_amd.Detector(dim, W, S, R, K, sched_kind=sched, index_kind=index, flags=flags)
    ds = rtod_amd.DeviceStream(xs[: W + (warmup + steps) * S])
    det.prime(xs[:W])
    if warmup:
        det.slides_device(ds.at(W), warmup)
    st0 = det.stats()
    t0 = time.perf_counter()
    if flags:
        det.reset_kernel_times()
    lists = det.slides_device(ds.at(W + warmup * S), steps)
    dt = time.perf_counter() - t0
    res = {"us": dt / steps * 1e6, "lists": lists, "stats": det.stats(), "stats_warm": st0,
           "kernels": det.kernel_times()}
    det.close()
    ds.close()
    return res


def cpu_baseline(xs, W, S, R, K, slides):
    import oracle

    threads = min(os.cpu_count() or 1, 16)
    _, us_mt = oracle.run_stream(xs[: W + slides * S], W, S, slides, R, K, "grid",
                                 nthreads=threads, with_time=True)
    _, us_st = oracle.run_stream(xs[: W + slides * S], W, S, slides, R, K, "grid",
                                 nthreads=1, with_time=True)
    return us_mt, us_st, threads


def calib_copy(mb: int = 64, reps: int = 5) -> None:
    """A plain device-to-device copy of known size (HIP's copy kernel), for
    scaling the FETCH_SIZE / WRITE_SIZE counters of a rocprofv3 --pmc run."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    n = mb * (1 << 20)
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"[calib] copied {mb} MiB x {reps}: reads {mb * 1024} KiB, writes {mb * 1024} KiB "
        f"per copy dispatch")


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, scaled by a plain copy), or None."""
    path = os.path.join(ROOT, "profiles", f"pmc_traffic_{workload}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("kernel=") and f"k_{kernel}<" in line:
            for tok in line.split():
                if tok.startswith("traffic_bytes="):
                    return int(tok.split("=")[1])
    return None


def index_comparison(names, steps):
    """Per-slide time of the cell-stencil index and of the LBVH index on the
    same streams (the north star picks whichever measures faster)."""
    import rtod_amd

    out = {}
    for name in names:
        dim, N, W, S, R, K, gen, seed = WORKLOADS[name]
        T = min(10000, (N - W) // S)
        xs = stream(name, T)
        n = min(steps, T - 10)
        st = run_gpu(xs, W, S, R, K, 10, n, index=rtod_amd.INDEX_CELL_STENCIL)
        lb = run_gpu(xs, W, S, R, K, 10, n, index=rtod_amd.INDEX_LBVH)
        out[name] = {"cell_stencil_us": round(st["lib_ms"] * 1e3, 3),
                     "lbvh_us": round(lb["lib_ms"] * 1e3, 3),
                     "faster": "cell_stencil" if st["lib_ms"] <= lb["lib_ms"] else "lbvh"}
        log(f"[bench] index comparison {name}: {out[name]}")
    return out


def learned_comparison(names, train_slides, steps):
    """Plain vs learned schedule on the same slides: the learned detector
    first trains online for `train_slides` slides (untimed warm-up, as the
    plain run's warm-up), then both are timed over `steps` slides."""
    import rtod_amd

    out = {}
    for name in names:
        dim, N, W, S, R, K, gen, seed = WORKLOADS[name]
        xs = stream(name, train_slides + steps)
        pl = run_gpu(xs, W, S, R, K, train_slides, steps, keep_lists=True)
        le = run_gpu(xs, W, S, R, K, train_slides, steps, sched=rtod_amd.SCHED_LEARNED,
                     keep_lists=True)
        mism = sum(0 if np.array_equal(a, b) else 1 for a, b in zip(pl["lists"], le["lists"]))
        pp = run_gpu_pipelined(xs, W, S, R, K, train_slides, steps)
        lp = run_gpu_pipelined(xs, W, S, R, K, train_slides, steps, sched=rtod_amd.SCHED_LEARNED)
        mism += sum(0 if np.array_equal(a, b) else 1 for a, b in zip(pl["lists"], lp["lists"]))
        s1, s0 = le["stats"], le["stats_warm"]
        pts = max(1, s1["predictor_points"] - s0["predictor_points"])
        # kernel durations of both schedules (separate event-timed passes)
        kp = run_gpu(xs, W, S, R, K, train_slides, min(steps, 60),
                     flags=rtod_amd.FLAG_PHASE_TIMES)["kernels"]
        kl = run_gpu(xs, W, S, R, K, train_slides, min(steps, 60),
                     flags=rtod_amd.FLAG_PHASE_TIMES, sched=rtod_amd.SCHED_LEARNED)["kernels"]
        ent = {"plain_us": round(pl["lib_ms"] * 1e3, 3),
               "learned_us": round(le["lib_ms"] * 1e3, 3),
               "speedup_vs_plain": round(pl["lib_ms"] / le["lib_ms"], 3),
               "plain_pipelined_us": round(pp["us"], 3),
               "learned_pipelined_us": round(lp["us"], 3),
               "pipelined_speedup_vs_plain": round(pp["us"] / lp["us"], 3),
               "mismatched_slides": mism,
               "miss_rate": round((s1["predictor_misses"] - s0["predictor_misses"]) / pts, 4),
               "fallback_slides": s1["fallbacks"] - s0["fallbacks"],
               "train_slides": train_slides,
               "train_steps": s1["train_steps"],
               "pair_tests_per_slide": {
                   "plain": round((pl["stats"]["pair_tests_total"] -
                                   pl["stats_warm"]["pair_tests_total"]) / steps),
                   "learned": round((s1["pair_tests_total"] - s0["pair_tests_total"]) / steps)},
               "kernels_ms_plain": {k: round(v[0], 5) for k, v in kp.items() if v[1] > 0},
               "kernels_ms_learned": {k: round(v[0], 5) for k, v in kl.items() if v[1] > 0}}
        out[name] = ent
        log(f"[bench] learned {name}: {ent}")
    return out


def _batch_run(descs, data, sched, warmup, steps, tune_slides):
    """Autotune, warm up (the learned schedule trains meanwhile), then time
    `steps` steps; per-step digests of all outlier lists for comparison."""
    import hashlib

    import rtod_amd

    cfg = [(d.dim, d.W, d.S, d.R, d.K) for d in descs]
    dss = [rtod_amd.DeviceStream(x) for x in data]
    b = rtod_amd.Batch(cfg, sched_kind=sched)
    b.bind_device([ds.ptr.value for ds in dss], [x.shape[0] for x in data])
    tuning = b.autotune(tune_slides)
    for _ in range(warmup):
        b.step_device()
    outs, digests, t_tot = 0, [], 0.0
    for _ in range(steps):
        t0 = time.perf_counter()
        lists = b.step_device()
        t_tot += time.perf_counter() - t0
        outs += sum(len(x) for x in lists)
        h = hashlib.sha256()
        for x in lists:
            h.update(x.tobytes())
            h.update(b"|")
        digests.append(h.hexdigest())
    st = [b.stream_stats(i) for i in range(len(descs))]
    b.close()
    for ds in dss:
        ds.close()
    return t_tot / steps, outs / steps, digests, tuning, st


def batched(n_streams, warmup, steps, tune_slides, cpu_slides, with_cpu, learned_warmup=40):
    """The batched workload: n_streams independent streams (dim ~ U{1,2,3},
    W log-uniform on [2k, 100k], S = 1-5% of W, K in {20, 50}, R set for a
    median neighbour count ~2K), resident in HBM, advanced together; plain,
    then the learned schedule on the same steps."""
    import rtod_amd
    from rtod_amd import synth

    slides = tune_slides + max(warmup, learned_warmup) + steps
    t0 = time.perf_counter()
    descs, data = synth.batch_workload(n_streams, slides=slides)
    log(f"[bench] batched: {n_streams} streams generated in {time.perf_counter() - t0:.1f} s, "
        f"sum W {sum(d.W for d in descs)}, sum S {sum(d.S for d in descs)}")
    dt, out_per_step, dig_p, tuning, _ = _batch_run(descs, data, rtod_amd.SCHED_PLAIN,
                                                    learned_warmup, steps, tune_slides)
    log(f"[bench] batched autotune: {tuning}")
    dt_l, _, dig_l, tuning_l, st_l = _batch_run(descs, data, rtod_amd.SCHED_LEARNED,
                                                learned_warmup, steps, tune_slides)
    pts = max(1, sum(s["predictor_points"] for s in st_l))
    learned = {"us_per_step": round(dt_l * 1e6, 2),
               "speedup_vs_plain": round(dt / dt_l, 3),
               "mismatched_steps": sum(1 for a, c in zip(dig_p, dig_l) if a != c),
               "miss_rate": round(sum(s["predictor_misses"] for s in st_l) / pts, 4),
               "fallback_stream_slides": int(sum(s["fallbacks"] for s in st_l)),
               "train_steps": int(max(s["train_steps"] for s in st_l)),
               "tuning": tuning_l}
    us = dt * 1e6
    bytes_step = sum(4 * d.dim * (d.W + d.S) for d in descs) + 4 * out_per_step
    achieved = bytes_step / dt / 1e9
    ent = {"workload": f"batch-{n_streams} streams: dim U{{1,2,3}}, W log-U[2k,100k], "
                       f"S 1-5% of W, K in {{20,50}}, R for median count ~2K",
           "streams": n_streams, "sum_W": int(sum(d.W for d in descs)),
           "sum_S": int(sum(d.S for d in descs)),
           "us_per_step": round(us, 2),
           "new_points_per_s": round(sum(d.S for d in descs) / dt),
           "outliers_per_step": round(out_per_step, 1),
           "tuning": tuning,
           "learned": learned,
           "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                        "algorithmic_bytes": int(bytes_step)}}
    if with_cpu:
        ent["cpu_baseline"] = batched_cpu(descs, data, cpu_slides)
        ent["speedup_vs_cpu"] = round(ent["cpu_baseline"]["value"] / us, 2)
    log(f"[bench] batched: {ent}")
    return ent


def batched_cpu(descs, data, slides):
    """Grid oracle over the same streams, one stream per task on a pool of
    host threads (ctypes releases the GIL); wall time of `slides` steps."""
    import oracle
    from concurrent.futures import ThreadPoolExecutor

    threads = min(os.cpu_count() or 1, 16)
    grids = [oracle.Grid(d.dim, d.W, d.S, d.R, d.K) for d in descs]
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(lambda i: grids[i].prime(data[i][:descs[i].W]), range(len(descs))))
        t0 = time.perf_counter()
        for t in range(1, slides + 1):
            list(ex.map(lambda i: grids[i].slide(
                data[i][descs[i].W + (t - 1) * descs[i].S: descs[i].W + t * descs[i].S]),
                range(len(descs))))
        dt = (time.perf_counter() - t0) / slides
    return {"value": round(dt * 1e6, 1), "unit": "us/step", "cores": threads, "kind": "port",
            "sample": f"grid oracle (C), every stream's first {slides} slides, streams spread "
                      f"over a {threads}-thread pool, one barrier per step"}


def other_workloads(names, steps, cpu_slides, with_cpu):
    out = {}
    for name in names:
        dim, N, W, S, R, K, gen, seed = WORKLOADS[name]
        T = min(10000, (N - W) // S)
        xs = stream(name, T)
        n = min(steps, T - 10)
        g = run_gpu(xs, W, S, R, K, 10, n, keep_lists=True)
        gp = run_gpu_pipelined(xs, W, S, R, K, 10, n)
        ent = {"gpu_us_per_slide": round(gp["us"], 3),
               "sequential_us_per_slide": round(g["lib_ms"] * 1e3, 3),
               "pipelined_mismatched_slides": sum(
                   0 if np.array_equal(a, b) else 1 for a, b in zip(g["lists"], gp["lists"])),
               "outliers_per_slide": g["outliers_mean"],
               "candidates_per_slide": g["stats"]["candidates_last"]}
        if with_cpu:
            ns = min(cpu_slides, T)
            us_mt, us_st, threads = cpu_baseline(xs, W, S, R, K, ns)
            ent["cpu_us_per_slide"] = round(us_mt, 2)
            ent["cpu_threads"] = threads
            ent["cpu_us_per_slide_1thread"] = round(us_st, 2)
            ent["speedup_vs_cpu"] = round(min(us_mt, us_st) / ent["gpu_us_per_slide"], 2)
        out[name] = ent
        log(f"[bench] {name}: {ent}")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=160)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-slides", type=int, default=60)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-others", action="store_true",
                    help="skip the 1-D W=10k and 3-D workloads")
    ap.add_argument("--workload", default="walk1d-W100k-S5k", choices=sorted(WORKLOADS),
                    help="profiling runs only: the JSON line always names the workload run")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched-streams measurement")
    ap.add_argument("--batch-streams", type=int, default=256)
    ap.add_argument("--batched-only", action="store_true",
                    help="profiling runs: only the batched measurement, its JSON object")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-schedule comparison")
    ap.add_argument("--train-slides", type=int, default=200,
                    help="online-training warm-up of the learned schedule")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run only a plain 64 MiB device copy (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs a stream on a single MI355X "
              "(multi-GPU sharding is design-only); use --gpus 1", file=sys.stderr)
        return 2
    if args.calib_copy:
        calib_copy()
        return 0
    if args.batched_only:
        print(json.dumps(batched(args.batch_streams, min(args.warmup, 5), args.steps, 4, 3,
                                 not args.no_cpu)), flush=True)
        return 0

    dim, _, W, S, R, K, _, _ = WORKLOADS[args.workload]
    xs = stream(args.workload, args.warmup + args.steps)
    log(f"[bench] {args.workload}: {xs.shape[0]} points; warmup {args.warmup} "
        f"steps {args.steps}")

    import rtod_amd

    plain = run_gpu(xs, W, S, R, K, args.warmup, args.steps, keep_lists=True)
    log(f"[bench] plain, inputs in HBM, one slide per call: {plain['lib_ms'] * 1e3:.1f} us/slide "
        f"(outer {plain['outer_ms'] * 1e3:.1f})")
    pipe = run_gpu_pipelined(xs, W, S, R, K, args.warmup, args.steps)
    pipe_mism = sum(0 if np.array_equal(a, b) else 1 for a, b in zip(plain["lists"], pipe["lists"]))
    log(f"[bench] plain, inputs in HBM, pipelined slides: {pipe['us']:.1f} us/slide "
        f"({pipe_mism} slides differ from the one-per-call run)")
    pcie = run_gpu(xs, W, S, R, K, args.warmup, args.steps, device_input=False)
    log(f"[bench] plain, inputs from host memory (PCIe inside the call): "
        f"{pcie['lib_ms'] * 1e3:.1f} us/slide")
    # per-kernel device durations: separate passes with HIP events around each
    # kernel, of the pipelined launches (the measured `value`) and of the
    # one-slide-per-call ones
    timed = run_gpu_pipelined(xs, W, S, R, K, min(args.warmup, 10), min(args.steps, 60),
                              flags=rtod_amd.FLAG_PHASE_TIMES)
    kern = {k: v[0] for k, v in timed["kernels"].items() if v[1] > 0}
    timed_seq = run_gpu(xs, W, S, R, K, min(args.warmup, 10), min(args.steps, 60),
                        flags=rtod_amd.FLAG_PHASE_TIMES)
    kern_seq = {k: v[0] for k, v in timed_seq["kernels"].items() if v[1] > 0}
    dom = max(kern, key=kern.get)
    dom_ms = kern[dom]
    n_out = plain["outliers_mean"]
    n_cand = plain["stats"]["candidates_last"]
    # algorithmic bytes per launch: 4*d bytes per point coordinate read or
    # written once (the per-point figure of the slide, B = 4d(W+S) + 4|O|)
    alg = {"ingest_sort": 8 * dim * S,                  # new slide in, ring out
           "merge_pass": 8 * dim * S,
           "merge_window": 8 * dim * W,                 # window in, window out
           "merge_window_ingest": 8 * dim * (W + S),    # + the next slide's tile sorts
           "classify": 4 * dim * W,                     # one pass over the window
           "detect": 4 * dim * n_cand + 4 * n_out,      # candidates in, list out
           "detect_runs": 4 * dim * n_cand + 4 * n_out + 8 * dim * S,  # + next run merge
           # two launches per slide (learned schedule): classify; detect +
           # the previous slide's list + the next slide's window merge + the
           # run merge two slides ahead + the tile sorts three ahead
           "classify_stage": 4 * dim * W,
           "detect_stage": 4 * dim * n_cand + 4 * n_out + 8 * dim * W + 16 * dim * S,
           # one launch per slide (plain schedule): classify + all of the above
           "slide_stage": 4 * dim * W + 4 * dim * n_cand + 4 * n_out + 8 * dim * W
                          + 16 * dim * S}
    b_dom = alg.get(dom, 4 * dim * (W + S))
    achieved = b_dom / (dom_ms * 1e-3) / 1e9
    b_slide = 4 * dim * (W + S) + 4 * n_out
    log(f"[bench] kernels (ms): " + ", ".join(f"{k}={v:.4f}" for k, v in kern.items()))
    traffic = measured_traffic(args.workload, dom)

    others = None
    if not args.no_others and args.workload == "walk1d-W100k-S5k":
        others = other_workloads(["walk1d-W10k-S500", "gmm3d-W10k-S500", "gmm3d-W100k-S5k"],
                                 120, args.cpu_slides, not args.no_cpu)

    indexes = None
    if not args.no_others and args.workload == "walk1d-W100k-S5k":
        indexes = index_comparison(["walk1d-W100k-S5k", "gmm3d-W10k-S500"], 60)

    batch = None
    if not args.no_batched and args.workload == "walk1d-W100k-S5k":
        batch = batched(args.batch_streams, 5, 40, 4, 3, not args.no_cpu)

    learned = None
    if not args.no_learned:
        names = [args.workload]
        if not args.no_others and args.workload == "walk1d-W100k-S5k":
            names += ["walk1d-W10k-S500", "gmm3d-W10k-S500", "gmm3d-W100k-S5k"]
        learned = learned_comparison(names, args.train_slides, args.steps)

    cpu = None
    if not args.no_cpu:
        us_mt, us_st, threads = cpu_baseline(xs, W, S, R, K, args.cpu_slides)
        best_mt = us_mt < us_st
        cpu = {"value": round(min(us_mt, us_st), 2), "unit": "us/slide",
               "cores": threads if best_mt else 1, "kind": "port",
               "sample": f"grid oracle (incremental per-cell FIFOs, C), slides "
                         f"1..{args.cpu_slides} of the same stream, the faster of a "
                         f"{threads}-thread pool ({us_mt:.1f} us/slide; range counts split "
                         f"over the pool only when a slide has >= 256 candidates) and 1 "
                         f"thread ({us_st:.1f} us/slide)"}
        log(f"[bench] cpu grid oracle: {us_mt:.1f} us/slide ({threads} thr), {us_st:.1f} (1 thr)")

    value = pipe["us"]
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "us/slide",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(value * 1e-3, 6),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded 1-D random walk (start 10, N(0,0.05) steps, 0.2% spikes "
                "+-U(2,5)); 3-D: mixture of 8 drifting anisotropic Gaussians + 0.5% noise",
        "config": {"workload": f"{args.workload} W={W} S={S} R={R} K={K} dim={dim} "
                               "(plain schedule, cell-stencil index; consecutive slides "
                               "pipelined on one stream, one launch per slide)",
                   "window": W, "slide": S, "R": R, "K": K, "dim": dim},
        "sequential_us_per_slide": round(plain["lib_ms"] * 1e3, 3),
        "pipelined_mismatched_slides": pipe_mism,
        "outer_us_per_slide": round(plain["outer_ms"] * 1e3, 3),
        "pcie_inclusive_us_per_slide": round(pcie["lib_ms"] * 1e3, 3),
        "outliers_per_slide": plain["outliers_mean"],
        "candidates_per_slide": plain["stats"]["candidates_last"],
        "kernels_ms": {k: round(v, 5) for k, v in kern.items()},
        "kernels_ms_one_slide_per_call": {k: round(v, 5) for k, v in kern_seq.items()},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 3),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic,
                     "algorithmic_bytes": int(b_dom), "kernel_avg_ms": round(dom_ms, 5),
                     "slide_bytes": int(b_slide),
                     "slide_frac": round(b_slide / (value * 1e-6) / 1e9 / HBM_PEAK_GBS, 6)},
        "cpu_baseline": cpu,
    }
    if cpu:
        line["speedup_vs_cpu"] = round(cpu["value"] / value, 2)
    if others:
        line["other_workloads"] = others
    if indexes:
        line["index_comparison"] = indexes
    if learned:
        line["learned"] = learned
    if batch:
        line["batched"] = batch
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
