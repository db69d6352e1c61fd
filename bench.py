#!/usr/bin/env python3
"""Per-slide time of the MI355X RTOD path on the 1-D random-walk stream
W=100k, S=5k, R=0.45, K=50 (1,048,572 points); one JSON line on stdout.

A step is one slide: S new points in host memory in, the window's exact
outlier list out in host memory.

  value    --steps consecutive slides through rtod_slides (pipelined, plain
           schedule, the index AUTO picks: the slide-run index in 1-D, the
           cell stencil in 2-D/3-D), timed on the host from the call's entry
           to its return: every slide's points are copied from host memory
           into pinned staging inside the timed call (the reference's
           per-slide H2D copy is inside its timed loop, optixScan.cpp:1051-
           1058) and every list is in host memory when it returns.  --warmup
           slides go through the same call first, untimed.  Before it, one
           untimed run of the same path on a context of its own warms the
           process (a process's first such run measured 0.1-0.3 us per slide
           slower than the next, fresh context or not).
  latency_us_per_slide
           the same slides one rtod_slide call each (new slide in host
           memory, list in host memory when the call returns), the mean of
           the library's wall time per call, with the reference's five Timer
           buckets (timer.h:16-24) from the kernels' device stamps.
  pipelined_hbm_us_per_slide, pipelined_pinned_us_per_slide,
  latency_pinned_us_per_slide
           the same pipelined slides already in HBM (rtod_slides_device),
           and in pinned host memory from rtod_host_alloc, which the
           kernels read in place over PCIe (no staging copy), pipelined and
           one rtod_slide call each.

Extra objects on the line:
  roofline      the dominant kernel of the measured path (the pipelined
                run's launch: a batch detected and the next sorted per launch
                on the slide-run index (32, then 128 slides a batch), one
                slide per launch on the cell
                stencil), timed with HIP events on the library's stream in a
                separate launch-by-launch pass from HBM (on the slide-run
                index a first batch of 32, then five full batches of 128),
                against the 8 TB/s HBM peak;
                algorithmic bytes per launch = the slides it detects times
                the slide's B = 4*d*W + 4*d*S + 4*|O| (window coordinates
                read once, new slide written, outlier ids written); traffic =
                HBM bytes per launch from the committed rocprofv3 counter
                summary
  cpu_baseline  the grid oracle (same algorithm, C, pthreads) on the host's
                cores: all of them and one thread, the faster reported, with
                the CPU model; and the number of points whose outlier status
                differs between this build's FP32 d2 <= R*R and the
                reference's FP64 strict test (optixScan.cu:170-189) over
                every slide of the stream
  sweeps        the reference's R, K and S sweeps around this default
                (script/run.py:76-218)
  other_workloads, index_comparison, ablations, learned, batched
                the 1-D W=10k and 3-D configurations, slide-run index vs
                cell stencil vs LBVH (1-D) and cell stencil vs LBVH (3-D),
                the reference's OPTIMIZATION 0/1/2 and refit-vs-rebuild
                modes and its ray-count runs, the learned work scheduler
                against the plain one, and 256 batched streams
The line carries "valid": false, and the process exits 1, if any pipelined,
learned or batched outlier list differs from the plain one-per-call list.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "μs per slide (W=100k,S=5k,K=50) plain vs learned-scheduled; exact outliers"
BUCKETS = ("copy_h2d", "prepare_cell", "build_index", "detect_outlier", "copy_d2h", "launch",
           "total")


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def workloads():
    from rtod_amd import synth

    return synth.WORKLOADS


def stream(name: str, slides: int) -> np.ndarray:
    """The measured stream; runs longer than N extend the same generator."""
    wl = workloads()[name]
    return wl.data(wl.W + slides * wl.S)


def same(a, b) -> int:
    """Slides whose lists differ."""
    return sum(0 if np.array_equal(x, y) else 1 for x, y in zip(a, b)) + abs(len(a) - len(b))


# --------------------------------------------------------------------------
# GPU legs
# --------------------------------------------------------------------------
def host_pipelined(xs, W, S, R, K, warmup, steps, sched=0, index=0, flags=0):
    """rtod_slides from host memory: `warmup` slides in one call, then
    `steps` slides in one timed call."""
    import rtod_amd

    dim = xs.shape[1]
    with rtod_amd.Detector(dim, W, S, R, K, sched_kind=sched, index_kind=index,
                           flags=flags) as det:
        det.prime(xs[:W])
        if warmup:
            det.slides(xs[W: W + warmup * S], warmup)
        st0 = det.stats()
        rows = np.ascontiguousarray(xs[W + warmup * S: W + (warmup + steps) * S])
        out = det._out_arrays(steps)
        t0 = time.perf_counter()
        det.slides_raw(rows, steps, out)  # the C-ABI call: slides in host memory to lists in host memory
        dt = time.perf_counter() - t0
        lists = det.split(*out)
        st1 = det.stats()
    return {"us": dt / steps * 1e6, "lists": lists, "stats": st1, "stats_warm": st0}


def pinned_pipelined(xs, W, S, R, K, warmup, steps, index=0):
    """rtod_slides from a buffer of rtod_host_alloc: the kernels read the
    slides in place over PCIe, with no staging copy (the rows are copied into
    the buffer before the timed call).  The buffer is used as a receive
    buffer is, again and again: a first, untimed run on a context of its own
    reads it once (the device's first pass over fresh pinned pages measured
    0.1-0.2 us per slide slower)."""
    import rtod_amd

    dim = xs.shape[1]
    buf = rtod_amd.pinned_empty(steps * S, dim)
    buf[:] = xs[W + warmup * S: W + (warmup + steps) * S]
    for _ in range(2):  # the second run is the one reported
        with rtod_amd.Detector(dim, W, S, R, K, index_kind=index) as det:
            det.prime(xs[:W])
            if warmup:
                det.slides(xs[W: W + warmup * S], warmup)
            out = det._out_arrays(steps)
            t0 = time.perf_counter()
            det.slides_raw(buf, steps, out)  # the C-ABI call: slides read in place
            dt = time.perf_counter() - t0
            lists = det.split(*out)
    del buf
    return {"us": dt / steps * 1e6, "lists": lists}


def device_pipelined(xs, W, S, R, K, warmup, steps, sched=0, index=0, flags=0):
    """rtod_slides_device: the stream uploaded to HBM first (untimed)."""
    import rtod_amd

    dim = xs.shape[1]
    with rtod_amd.DeviceStream(xs[: W + (warmup + steps) * S]) as ds, \
            rtod_amd.Detector(dim, W, S, R, K, sched_kind=sched, index_kind=index,
                              flags=flags) as det:
        det.prime(xs[:W])
        if warmup:
            det.slides_device(ds.at(W), warmup)
        if flags:
            det.reset_kernel_times()
        out = det._out_arrays(steps)
        t0 = time.perf_counter()
        det.slides_device_raw(ds.at(W + warmup * S), steps, out)
        dt = time.perf_counter() - t0
        lists = det.split(*out)
        kern = det.kernel_times()
    return {"us": dt / steps * 1e6, "lists": lists, "kernels": kern}


def single_calls(xs, W, S, R, K, warmup, steps, sched=0, index=0, flags=0, pinned=False):
    """One rtod_slide call per slide (host input): the library's wall time
    per call, its buckets and kernel durations (device stamps).  pinned: the
    slides in a buffer from rtod_host_alloc (read in place, no staging copy)."""
    import rtod_amd

    dim = xs.shape[1]
    src, base = xs, W
    if pinned:
        src = rtod_amd.pinned_empty((warmup + steps) * S, dim)
        src[:] = xs[W: W + (warmup + steps) * S]
        base = 0
    with rtod_amd.Detector(dim, W, S, R, K, sched_kind=sched, index_kind=index,
                           flags=flags) as det:
        det.prime(xs[:W])
        lists, bk, stg, outer = [], [], [], 0.0
        st0 = None
        for t in range(1, warmup + steps + 1):
            row = base + (t - 1) * S
            t0 = time.perf_counter()
            o = det.slide(src[row: row + S])
            dt = time.perf_counter() - t0
            if t == warmup:
                st0 = det.stats()
            if t > warmup:
                tm = det.times
                bk.append([getattr(tm, b) for b in BUCKETS])
                stg.append(list(det.stage_times().values()))
                lists.append(o)
                outer += dt
        st1 = det.stats()
        kern = det.kernel_times()
    bk = np.array(bk).mean(0)
    return {"us": bk[-1] * 1e3, "outer_us": outer / steps * 1e6, "lists": lists,
            "buckets_ms": {b: round(float(v), 5) for b, v in zip(BUCKETS, bk)},
            "stages_ms": {n: round(float(v), 5)
                          for n, v in zip(rtod_amd.Detector.STAGES, np.array(stg).mean(0))},
            "stats": st1, "stats_warm": st0, "kernels": kern}


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, scaled by a plain copy), or None."""
    path = os.path.join(ROOT, "profiles", f"pmc_traffic_{workload}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        name = line.split()[0][len("kernel="):] if line.startswith("kernel=") else ""
        if name == f"rtod::k_{kernel}" or name.startswith(f"rtod::k_{kernel}<"):
            for tok in line.split():
                if tok.startswith("traffic_bytes="):
                    return int(tok.split("=")[1])
    return None


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


# --------------------------------------------------------------------------
# CPU baseline (the oracle: test infrastructure, timed here as the baseline)
# --------------------------------------------------------------------------
def host_threads() -> tuple[int, dict]:
    """The host CPUs this process may use: its affinity mask, capped by the
    cgroup's CPU quota and by OMP_NUM_THREADS when set (a machine's share
    for this job), with what each said."""
    info = {"cpu_count": os.cpu_count()}
    n = len(os.sched_getaffinity(0))
    info["affinity"] = n
    try:
        q, p = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            info["cgroup_quota"] = round(int(q) / int(p), 2)
            n = min(n, max(1, int(int(q) / int(p))))
    except (OSError, ValueError):
        pass
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        info["omp_num_threads"] = int(env)
        n = min(n, int(env))
    return n, info


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(xs, W, S, R, K, slides, threads_all):
    """Grid oracle, slides 1..slides: all host threads and one thread."""
    import oracle

    _, us_mt = oracle.run_stream(xs[: W + slides * S], W, S, slides, R, K, "grid",
                                 nthreads=threads_all, with_time=True)
    _, us_st = oracle.run_stream(xs[: W + slides * S], W, S, slides, R, K, "grid",
                                 nthreads=1, with_time=True)
    best_mt = us_mt < us_st
    return {"value": round(min(us_mt, us_st), 2), "unit": "us/slide",
            "cores": threads_all if best_mt else 1, "kind": "port",
            "sample": f"grid oracle (incremental per-cell FIFOs, C), slides 1..{slides} of the "
                      f"same stream: {threads_all} threads {us_mt:.1f} us/slide, 1 thread "
                      f"{us_st:.1f} us/slide, the faster reported (range counts split over "
                      f"the threads only when a slide has >= 256 candidates)",
            "us_all_threads": round(us_mt, 2), "us_1_thread": round(us_st, 2),
            "threads_all": threads_all, "host_cpus": host_threads()[1],
            "cpu_model": cpu_model()}


def fp64_disagreements(xs, W, S, T, R, K, threads):
    """Points whose outlier status differs between the FP32 contract and the
    reference's FP64 strict rule, every slide (oracle, CPU)."""
    import oracle

    t0 = time.perf_counter()
    tot, per = oracle.semantics_diff_stream(xs, W, S, T, R, K, nthreads=threads)
    return {"slides": int(T), "points_status_differs": tot["status"],
            "slides_with_a_difference": int((per > 0).sum()),
            "points_count_differs": tot["counts"], "pairs_differ": tot["pairs"],
            "seconds": round(time.perf_counter() - t0, 1)}


# --------------------------------------------------------------------------
# workloads beside the headline
# --------------------------------------------------------------------------
def one_workload(name, warmup, steps, cpu_slides, with_cpu, threads, R=None, K=None, S=None):
    wl = workloads()[name]
    W = wl.W
    S = S or wl.S
    R = R or wl.R
    K = K or wl.K
    T = min(10000, (wl.N - W) // S)
    n = min(steps, T - warmup)
    xs = wl.data(W + T * S)
    hp = host_pipelined(xs, W, S, R, K, warmup, n)
    sc = single_calls(xs, W, S, R, K, warmup, n)
    ent = {"config": {"dim": wl.dim, "window": W, "slide": S, "R": R, "K": K, "slides": n},
           "gpu_us_per_slide": round(hp["us"], 3),
           "latency_us_per_slide": round(sc["us"], 3),
           "breakdown_ms": sc["buckets_ms"],
           "mismatched_slides": same(hp["lists"], sc["lists"]),
           "outliers_per_slide": round(float(np.mean([len(x) for x in sc["lists"]])), 2),
           "candidates_per_slide": round((sc["stats"]["candidates_total"] -
                                          sc["stats_warm"]["candidates_total"]) / n, 1)}
    if with_cpu:
        cb = cpu_baseline(xs, W, S, R, K, min(cpu_slides, T), threads)
        ent["cpu_us_per_slide"] = cb["value"]
        ent["cpu_cores"] = cb["cores"]
        ent["cpu_us_1_thread"] = cb["us_1_thread"]
        ent["cpu_us_all_threads"] = cb["us_all_threads"]
        ent["speedup_vs_cpu"] = round(cb["value"] / hp["us"], 2)
        ent["latency_speedup_vs_cpu"] = round(cb["value"] / sc["us"], 2)
    return ent


def sweeps(warmup, steps, cpu_slides, with_cpu, threads):
    out = {}
    for what, vals in (("R", (0.15, 0.3, 0.6, 0.75)), ("K", (30, 70, 100)),
                       ("S", (1000, 10000))):
        for v in vals:
            kw = {what: v}
            key = f"{what}={v}"
            out[key] = one_workload("walk1d-W100k-S5k", warmup, steps, cpu_slides, with_cpu,
                                    threads, **kw)
            log(f"[bench] sweep {key}: {out[key]}")
    return out


def index_comparison(names, warmup, steps):
    """Per-slide time of each index on the same streams, one slide per call
    and host-fed pipelined: the slide-run index (1-D only), the cell stencil
    and the LBVH (AUTO takes the fastest measured)."""
    import rtod_amd

    kinds = {"slide_runs": rtod_amd.INDEX_SLIDE_RUNS,
             "cell_stencil": rtod_amd.INDEX_CELL_STENCIL, "lbvh": rtod_amd.INDEX_LBVH}
    out = {}
    for name in names:
        wl = workloads()[name]
        xs = stream(name, warmup + steps)
        ent, ref = {}, None
        for k, ix in kinds.items():
            if k == "slide_runs" and wl.dim != 1:
                continue
            sc = single_calls(xs, wl.W, wl.S, wl.R, wl.K, warmup, steps, index=ix)
            hp = host_pipelined(xs, wl.W, wl.S, wl.R, wl.K, warmup, steps, index=ix)
            ref = sc["lists"] if ref is None else ref
            ent[f"{k}_us"] = round(sc["us"], 3)
            ent[f"{k}_pipelined_us"] = round(hp["us"], 3)
            ent[f"{k}_stages_ms"] = sc["stages_ms"]
            ent.setdefault("mismatched_slides", 0)
            ent["mismatched_slides"] += same(sc["lists"], ref) + same(hp["lists"], ref)
        ent["faster"] = min((k for k in kinds if f"{k}_us" in ent),
                            key=lambda k: ent[f"{k}_pipelined_us"])
        out[name] = ent
        log(f"[bench] index comparison {name}: {out[name]}")
    return out


def ablations(names, warmup, steps):
    """The reference's OPTIMIZATION 0/1/2 and UPDATE_GAS_TYPE comparison
    (its Figure 15 runs, script/run.py:254-270) on the same streams: one
    rtod_slide call per slide in each mode, every list checked against the
    default path's; and its ray-count experiment (script/run.py:237-251): no
    grid filter, an LBVH over the window, launch_ray_num rays."""
    import rtod_amd as RA

    modes = {
        "default": dict(),
        "stencil_default": dict(index_kind=RA.INDEX_CELL_STENCIL),
        "stencil_no_filter": dict(ablation=RA.ABL_NO_FILTER),
        "lbvh_default_opt2": dict(index_kind=RA.INDEX_LBVH),
        "lbvh_filter_only_opt1": dict(index_kind=RA.INDEX_LBVH, ablation=RA.ABL_FILTER_ONLY),
        "lbvh_filter_only_opt1_refit": dict(index_kind=RA.INDEX_LBVH,
                                            ablation=RA.ABL_FILTER_ONLY, refit=True),
        "lbvh_no_filter_opt0": dict(index_kind=RA.INDEX_LBVH, ablation=RA.ABL_NO_FILTER),
        "lbvh_no_filter_opt0_refit": dict(index_kind=RA.INDEX_LBVH, ablation=RA.ABL_NO_FILTER,
                                          refit=True),
    }
    out = {}
    for name in names:
        wl = workloads()[name]
        xs = stream(name, warmup + steps)
        ref, ent = None, {}
        for m, kw in modes.items():
            r = _single_kw(xs, wl.W, wl.S, wl.R, wl.K, warmup, steps, kw)
            if ref is None:
                ref = r["lists"]
            ent[m] = {"us_per_slide": round(r["us"], 3), "stages_ms": r["stages_ms"],
                      "pair_tests_per_slide": round(r["tests"] / steps),
                      "candidates_per_slide": round(r["cands"] / steps, 1),
                      "mismatched_slides": same(r["lists"], ref)}
        out[name] = ent
        log(f"[bench] ablations {name}: {ent}")
    wl = workloads()["walk1d-W100k-S5k"]
    xs = stream(wl.name, warmup + steps)
    rays = {}
    for L in (1000, 10000, 100000):
        kw = dict(index_kind=RA.INDEX_LBVH, ablation=RA.ABL_NO_FILTER, launch_ray_num=L)
        r = _single_kw(xs, wl.W, wl.S, wl.R, wl.K, warmup, steps, kw)
        rays[str(L)] = {"us_per_slide": round(r["us"], 3),
                        "detect_us": round(r["stages_ms"]["detect"] * 1e3, 3),
                        "rays_per_s_detect": round(L / (r["stages_ms"]["detect"] * 1e-3))}
    out["ray_count_walk1d-W100k-S5k"] = rays
    log(f"[bench] ray counts: {rays}")
    return out


def _single_kw(xs, W, S, R, K, warmup, steps, kw):
    """single_calls with detector options; counters of the timed slides."""
    import rtod_amd

    dim = xs.shape[1]
    with rtod_amd.Detector(dim, W, S, R, K, **kw) as det:
        det.prime(xs[:W])
        lists, tot, stg = [], [], []
        for t in range(1, warmup + steps + 1):
            if t == warmup + 1:
                st0 = det.stats()
            o = det.slide(xs[W + (t - 1) * S: W + t * S])
            if t > warmup:
                lists.append(o)
                tot.append(det.times.total)
                stg.append(list(det.stage_times().values()))
        st1 = det.stats()
    return {"us": float(np.mean(tot)) * 1e3, "lists": lists,
            "stages_ms": {n: round(float(v), 5)
                          for n, v in zip(rtod_amd.Detector.STAGES, np.array(stg).mean(0))},
            "tests": st1["pair_tests_total"] - st0["pair_tests_total"],
            "cands": st1["candidates_total"] - st0["candidates_total"]}


def learned_comparison(names, train_slides, steps):
    """Plain vs learned schedule on the same slides: each detector first runs
    `train_slides` slides (untimed warm-up; the learned one trains online
    meanwhile), then `steps` timed slides, host-fed pipelined and one per
    call.  The learned schedule as shipped applies its cost rule (plain
    kernels where prediction cannot shorten a slide); `learned_kernels_*`
    force the learned kernels on every slide (RTOD_FLAG_LEARN_ALWAYS), which
    shows what inference costs and what the ordering saves."""
    import rtod_amd

    L, ALW = rtod_amd.SCHED_LEARNED, rtod_amd.FLAG_LEARN_ALWAYS
    out = {}
    for name in names:
        wl = workloads()[name]
        xs = stream(name, train_slides + steps)
        W, S, R, K = wl.W, wl.S, wl.R, wl.K
        pl = single_calls(xs, W, S, R, K, train_slides, steps)
        le = single_calls(xs, W, S, R, K, train_slides, steps, sched=L)
        lk = single_calls(xs, W, S, R, K, train_slides, steps, sched=L, flags=ALW)
        pp = host_pipelined(xs, W, S, R, K, train_slides, steps)
        lp = host_pipelined(xs, W, S, R, K, train_slides, steps, sched=L)
        lkp = host_pipelined(xs, W, S, R, K, train_slides, steps, sched=L, flags=ALW)
        mism = sum(same(pl["lists"], x["lists"]) for x in (le, lk, pp, lp, lkp))

        def delta(r, k):
            return r["stats"][k] - r["stats_warm"][k]

        pts = max(1, delta(lk, "predictor_points"))
        ent = {"plain_us": round(pl["us"], 3), "learned_us": round(le["us"], 3),
               "speedup_vs_plain": round(pl["us"] / le["us"], 3),
               "plain_pipelined_us": round(pp["us"], 3),
               "learned_pipelined_us": round(lp["us"], 3),
               "pipelined_speedup_vs_plain": round(pp["us"] / lp["us"], 3),
               "plain_kernel_slides": delta(le, "plain_kernel_slides"),
               "fallback_slides": delta(le, "fallbacks"),
               "learned_kernels_us": round(lk["us"], 3),
               "learned_kernels_speedup_vs_plain": round(pl["us"] / lk["us"], 3),
               "learned_kernels_pipelined_us": round(lkp["us"], 3),
               "learned_kernels_pipelined_speedup_vs_plain": round(pp["us"] / lkp["us"], 3),
               "learned_kernels_miss_rate": round(delta(lk, "predictor_misses") / pts, 4),
               "learned_kernels_fallback_slides": delta(lk, "fallbacks"),
               "mismatched_slides": mism,
               "train_slides": train_slides, "train_steps": lk["stats"]["train_steps"],
               "units_per_slide": round(delta(lk, "predictor_points") / steps),
               "pair_tests_per_slide": {
                   "plain": round(delta(pl, "pair_tests_total") / steps),
                   "learned": round(delta(le, "pair_tests_total") / steps),
                   "learned_kernels": round(delta(lk, "pair_tests_total") / steps)},
               "stages_ms_plain": pl["stages_ms"], "stages_ms_learned": le["stages_ms"],
               "stages_ms_learned_kernels": lk["stages_ms"]}
        out[name] = ent
        log(f"[bench] learned {name}: {ent}")
    return out


def _batch_run(descs, data, sched, warmup, steps, tune_slides):
    """Autotune, warm up (the learned schedule trains meanwhile), then time
    `steps` steps; per-step digests of all outlier lists for comparison."""
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


def batched(n_streams, steps, tune_slides, cpu_slides, with_cpu, threads, learned_warmup=40):
    """256 independent streams (dim ~ U{1,2,3}, W log-uniform on [2k, 100k],
    S = 1-5% of W, K in {20, 50}, R set for a median neighbour count ~2K),
    resident in HBM, advanced together; plain, then the learned schedule on
    the same steps."""
    import rtod_amd
    from rtod_amd import synth

    slides = tune_slides + learned_warmup + steps
    t0 = time.perf_counter()
    descs, data = synth.batch_workload(n_streams, slides=slides)
    log(f"[bench] batched: {n_streams} streams generated in {time.perf_counter() - t0:.1f} s, "
        f"sum W {sum(d.W for d in descs)}, sum S {sum(d.S for d in descs)}")
    dt, out_per_step, dig_p, tuning, _ = _batch_run(descs, data, rtod_amd.SCHED_PLAIN,
                                                    learned_warmup, steps, tune_slides)
    dt_l, _, dig_l, tuning_l, st_l = _batch_run(descs, data, rtod_amd.SCHED_LEARNED,
                                                learned_warmup, steps, tune_slides)
    pts = max(1, sum(s["predictor_points"] for s in st_l))
    learned = {"us_per_step": round(dt_l * 1e6, 2),
               "speedup_vs_plain": round(dt / dt_l, 3),
               "mismatched_steps": sum(1 for a, c in zip(dig_p, dig_l) if a != c),
               "miss_rate": round(sum(s["predictor_misses"] for s in st_l) / pts, 4),
               "fallback_stream_slides": int(sum(s["fallbacks"] for s in st_l)),
               "plain_kernel_stream_slides": int(sum(s["plain_kernel_slides"] for s in st_l)),
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
        ent["cpu_baseline"] = batched_cpu(descs, data, cpu_slides, threads)
        ent["speedup_vs_cpu"] = round(ent["cpu_baseline"]["value"] / us, 2)
    log(f"[bench] batched: {ent}")
    return ent


def batched_cpu(descs, data, slides, threads):
    """Grid oracle over the same streams, one stream per task on a pool of
    host threads (ctypes releases the GIL); wall time of `slides` steps."""
    import oracle
    from concurrent.futures import ThreadPoolExecutor

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
                      f"over a {threads}-thread pool, one barrier per step",
            "cpu_model": cpu_model()}


# --------------------------------------------------------------------------
def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=160)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-slides", type=int, default=60)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-others", action="store_true",
                    help="skip the other workloads, the sweeps and the index comparison")
    ap.add_argument("--workload", default="walk1d-W100k-S5k", choices=sorted(workloads()),
                    help="profiling runs only: the JSON line always names the workload run")
    ap.add_argument("--index", default="auto", choices=("auto", "runs", "stencil", "lbvh"),
                    help="profiling runs only: the index of the measured workload's legs")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batch-streams", type=int, default=256)
    ap.add_argument("--batched-only", action="store_true",
                    help="profiling runs: only the batched measurement, its JSON object")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--train-slides", type=int, default=200,
                    help="online-training warm-up of the learned schedule")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run only a plain 64 MiB device copy (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every GPU box of this build has one, and the "
              "multi-GPU sharding of a stream is design-only; use --gpus 1", file=sys.stderr)
        return 2
    if args.calib_copy:
        calib_copy()
        return 0
    threads = host_threads()[0]
    if args.batched_only:
        print(json.dumps(batched(args.batch_streams, args.steps, 4, 3, not args.no_cpu,
                                 threads)), flush=True)
        return 0

    import rtod_amd

    wl = workloads()[args.workload]
    dim, W, S, R, K = wl.dim, wl.W, wl.S, wl.R, wl.K
    xs = stream(args.workload, args.warmup + args.steps)
    log(f"[bench] {args.workload}: {xs.shape[0]} points; warmup {args.warmup} "
        f"steps {args.steps}; {rtod_amd.version()}")

    ix = {"auto": rtod_amd.INDEX_AUTO, "runs": rtod_amd.INDEX_SLIDE_RUNS,
          "stencil": rtod_amd.INDEX_CELL_STENCIL, "lbvh": rtod_amd.INDEX_LBVH}[args.index]
    # the measured value: host-fed pipelined slides, after one untimed run of
    # the same path on a context of its own (the process's first)
    host_pipelined(xs, W, S, R, K, args.warmup, args.steps, index=ix)
    hp = host_pipelined(xs, W, S, R, K, args.warmup, args.steps, index=ix)
    log(f"[bench] host-fed pipelined (rtod_slides): {hp['us']:.2f} us/slide")
    # one call per slide, host input: latency and the reference's buckets
    sc = single_calls(xs, W, S, R, K, args.warmup, args.steps, index=ix)
    log(f"[bench] one slide per call (rtod_slide): {sc['us']:.2f} us/slide "
        f"(outer {sc['outer_us']:.2f}); buckets {sc['buckets_ms']}; stages {sc['stages_ms']}")
    dp = device_pipelined(xs, W, S, R, K, args.warmup, args.steps, index=ix)
    log(f"[bench] HBM-fed pipelined (rtod_slides_device): {dp['us']:.2f} us/slide")
    scp = single_calls(xs, W, S, R, K, args.warmup, args.steps, index=ix, pinned=True)
    log(f"[bench] one slide per call from rtod_host_alloc memory: {scp['us']:.2f} us/slide")
    pp = pinned_pipelined(xs, W, S, R, K, args.warmup, args.steps, index=ix)
    log(f"[bench] host-fed pipelined from rtod_host_alloc memory (read in place): "
        f"{pp['us']:.2f} us/slide")
    mism = (same(hp["lists"], sc["lists"]) + same(dp["lists"], sc["lists"]) +
            same(pp["lists"], sc["lists"]) + same(scp["lists"], sc["lists"]))

    # the dominant kernel of the measured path, HIP events around each launch
    # a run in the steady state: on the slide-run index its first batch (32
    # slides) and five full batches from HBM, launch by launch
    n_roof = rtod_amd.RUN_FIRST_BATCH + 5 * rtod_amd.RUN_BATCH
    xs_roof = stream(args.workload, 10 + n_roof)
    timed = device_pipelined(xs_roof, W, S, R, K, 10, n_roof,
                             index=ix, flags=rtod_amd.FLAG_PHASE_TIMES)
    kern = {k: v[0] for k, v in timed["kernels"].items() if v[1] > 0}
    launches = {k: v[1] for k, v in timed["kernels"].items() if v[1] > 0}
    dom = max(kern, key=lambda k: kern[k] * launches[k])
    dom_ms = kern[dom]
    # slides one launch of the dominant kernel detects: on the slide-run
    # index the run's batches are RUN_FIRST_BATCH slides, then RUN_BATCH
    # each, and every launch but the last (run_detect, which only detects)
    # detects one batch and sorts the next (run_stage); one slide per launch
    # on the cell stencil
    if dom == "run_stage":
        rest = n_roof - rtod_amd.RUN_FIRST_BATCH
        last = (rest % rtod_amd.RUN_BATCH or rtod_amd.RUN_BATCH) if rest > 0 else n_roof
        per_launch = (n_roof - last) / launches[dom]
    else:
        per_launch = n_roof / launches[dom]
    n_out = float(np.mean([len(x) for x in sc["lists"]]))
    b_slide = 4 * dim * W + 4 * dim * S + 4 * n_out
    b_launch = b_slide * per_launch
    achieved = b_launch / (dom_ms * 1e-3) / 1e9
    traffic = measured_traffic(args.workload, dom)
    log(f"[bench] pipelined launch {dom}: {dom_ms * 1e3:.2f} us for {per_launch:.2f} slides; "
        f"B_slide {b_slide:.0f} B")
    index_used = {rtod_amd.INDEX_SLIDE_RUNS: "slide-run", rtod_amd.INDEX_CELL_STENCIL:
                  "cell-stencil", rtod_amd.INDEX_LBVH: "LBVH"}[hp["stats"]["index_kind"]]

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(xs, W, S, R, K, args.cpu_slides, threads)
        cpu["fp64_strict_disagreements"] = fp64_disagreements(wl.data(), W, S, wl.T, R, K,
                                                              threads)
        log(f"[bench] cpu: {cpu}")

    others = sweep = indexes = abl = None
    if not args.no_others and args.workload == "walk1d-W100k-S5k":
        others = {}
        for name in ("walk1d-W10k-S500", "gmm3d-W10k-S500", "gmm3d-W100k-S5k"):
            others[name] = one_workload(name, 10, 120, args.cpu_slides, not args.no_cpu, threads)
            if not args.no_cpu and name != "gmm3d-W100k-S5k":
                w2 = workloads()[name]
                others[name]["fp64_strict_disagreements"] = fp64_disagreements(
                    w2.data(), w2.W, w2.S, w2.T, w2.R, w2.K, threads)
            log(f"[bench] {name}: {others[name]}")
        sweep = sweeps(5, 40, 30, not args.no_cpu, threads)
        indexes = index_comparison(["walk1d-W100k-S5k", "gmm3d-W10k-S500"], 10, 40)
        abl = ablations(["walk1d-W100k-S5k", "gmm3d-W10k-S500"], 5, 20)

    batch = None
    if not args.no_batched and args.workload == "walk1d-W100k-S5k":
        batch = batched(args.batch_streams, 40, 4, 3, not args.no_cpu, threads)

    learned = None
    if not args.no_learned:
        names = [args.workload]
        if not args.no_others and args.workload == "walk1d-W100k-S5k":
            names += ["walk1d-W10k-S500", "gmm3d-W10k-S500", "gmm3d-W100k-S5k"]
        learned = learned_comparison(names, args.train_slides, min(args.steps, 120))

    invalid = mism
    invalid += sum(v["mismatched_slides"] for v in (others or {}).values())
    invalid += sum(v["mismatched_slides"] for v in (sweep or {}).values())
    invalid += sum(v["mismatched_slides"] for v in (indexes or {}).values())
    invalid += sum(m["mismatched_slides"] for k, v in (abl or {}).items()
                   if not k.startswith("ray_count") for m in v.values())
    invalid += sum(v["mismatched_slides"] for v in (learned or {}).values())
    if batch:
        invalid += batch["learned"]["mismatched_steps"]
    value = hp["us"]
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
        "config": {"workload": f"{args.workload} W={W} S={S} R={R} K={K} dim={dim} (plain "
                               f"schedule, {index_used} index; value: slides from pageable "
                               f"host memory, pipelined on one stream; roofline: the same "
                               f"kernel on slides resident in HBM, {per_launch:.2f} slides "
                               f"detected per launch)",
                   "window": W, "slide": S, "R": R, "K": K, "dim": dim},
        "valid": invalid == 0,
        "mismatched_slides": int(mism),
        "latency_us_per_slide": round(sc["us"], 3),
        "latency_outer_us_per_slide": round(sc["outer_us"], 3),
        "breakdown_ms": sc["buckets_ms"],
        "stages_ms": sc["stages_ms"],
        "pipelined_hbm_us_per_slide": round(dp["us"], 3),
        "pipelined_pinned_us_per_slide": round(pp["us"], 3),
        "latency_pinned_us_per_slide": round(scp["us"], 3),
        "outliers_per_slide": round(n_out, 2),
        "candidates_per_slide": round((sc["stats"]["candidates_total"] -
                                       sc["stats_warm"]["candidates_total"]) / args.steps, 1),
        "kernels_ms_pipelined": {k: round(v, 5) for k, v in kern.items()},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 3),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic,
                     "traffic_over_algorithmic": round(traffic / b_launch, 2) if traffic else None,
                     "algorithmic_bytes": int(b_launch), "slides_per_launch": round(per_launch, 2),
                     "algorithmic_bytes_per_slide": int(b_slide),
                     "kernel_avg_ms": round(dom_ms, 5),
                     "slide_frac": round(b_slide / (value * 1e-6) / 1e9 / HBM_PEAK_GBS, 6)},
        "cpu_baseline": cpu,
    }
    if cpu:
        line["speedup_vs_cpu"] = round(cpu["value"] / value, 2)
        line["latency_speedup_vs_cpu"] = round(cpu["value"] / sc["us"], 2)
    if others:
        line["other_workloads"] = others
    if sweep:
        line["sweeps"] = sweep
    if indexes:
        line["index_comparison"] = indexes
    if abl:
        line["ablations"] = abl
    if learned:
        line["learned"] = learned
    if batch:
        line["batched"] = batch
    print(json.dumps(line), flush=True)
    return 0 if invalid == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
