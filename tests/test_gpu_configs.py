"""GPU parity on the measured configurations, every slide, through the C-ABI.

The 1-D reference-shape stream (W=100k, S=5k) and the 3-D stream on all of
their slides, one slide per call and pipelined; the 1-D stream's R, K and S
sweeps (the reference's vary-R/K/S experiments, script/run.py:76-218) on all
of their slides; the learned schedule on the 1-D reference shape.  Oracles:
the grid oracle (same algorithm, every slide) and, in 1-D, the sorted-window
oracle (independent method) on the first slides.  Bar: bit-exact sets.
"""
import numpy as np
import pytest

import oracle as O
import rtod_amd as R

pytestmark = pytest.mark.gpu


def case(name):
    print(f"[case] {name}", flush=True)


def assert_same(got, want, name):
    assert len(got) == len(want), f"{name}: {len(got)} slides vs {len(want)}"
    for t, (g, w) in enumerate(zip(got, want), 1):
        if not np.array_equal(g, w):
            only_g = sorted(set(g.tolist()) - set(w.tolist()))[:8]
            only_w = sorted(set(w.tolist()) - set(g.tolist()))[:8]
            raise AssertionError(f"{name}: slide {t}: gpu-only {only_g} oracle-only {only_w}")


def pipelined(xs, W, S, T, Rr, K, **kw):
    dim = xs.shape[1]
    with R.DeviceStream(xs[: W + T * S]) as ds, R.Detector(dim, W, S, Rr, K, **kw) as det:
        det.prime(xs[:W])
        return det.slides_device(ds.at(W), T)


def both_paths(xs, W, S, T, Rr, K, name, **kw):
    want = O.run_stream(xs, W, S, T, Rr, K, "grid")
    assert_same(R.run_stream(xs, W, S, T, Rr, K, **kw), want, f"{name} one per call")
    assert_same(pipelined(xs, W, S, T, Rr, K, **kw), want, f"{name} pipelined")
    return want


INDEX_1D = {"runs": R.INDEX_SLIDE_RUNS, "stencil": R.INDEX_CELL_STENCIL}


@pytest.mark.parametrize("index", list(INDEX_1D))
def test_c2_every_slide(index):
    wl = O.C2
    case(f"{wl.name} all {wl.T} slides, {index}")
    xs = wl.data()
    want = both_paths(xs, wl.W, wl.S, wl.T, wl.R, wl.K, wl.name, index_kind=INDEX_1D[index])
    assert_same(want[:8], O.run_stream(xs, wl.W, wl.S, 8, wl.R, wl.K, "sorted1d"),
                "C2 grid vs sorted1d")


def test_c3_every_slide():
    wl = O.C3
    case(f"{wl.name} all {wl.T} slides")
    both_paths(wl.data(), wl.W, wl.S, wl.T, wl.R, wl.K, wl.name)


# the reference's sweeps around the 1-D default (R 0.45, K 50, S 5k)
SWEEP = [("R", 0.15), ("R", 0.3), ("R", 0.6), ("R", 0.75), ("K", 30), ("K", 70), ("K", 100),
         ("S", 1000), ("S", 10000)]


@pytest.mark.parametrize("index", list(INDEX_1D))
@pytest.mark.parametrize("what,val", SWEEP)
def test_c2_sweep_every_slide(what, val, index):
    wl = O.C2
    Rr, K, S = wl.R, wl.K, wl.S
    if what == "R":
        Rr = val
    elif what == "K":
        K = val
    else:
        S = val
    T = min(10000, (wl.N - wl.W) // S)
    case(f"{wl.name} sweep {what}={val}: W={wl.W} S={S} R={Rr} K={K}, {T} slides, {index}")
    xs = wl.data()
    want = both_paths(xs, wl.W, S, T, Rr, K, f"sweep {what}={val}", index_kind=INDEX_1D[index])
    assert_same(want[:4], O.run_stream(xs, wl.W, S, 4, Rr, K, "sorted1d"),
                f"sweep {what}={val}: grid vs sorted1d")


@pytest.mark.parametrize("flags", [0, R.FLAG_LEARN_ALWAYS], ids=["cost-rule", "learned-kernels"])
def test_learned_c2_every_slide(flags):
    """C2 with the learned schedule: as shipped (1-D, so the cost rule runs
    the plain kernels) and with the learned kernels forced on every slide."""
    wl = O.C2
    case(f"{wl.name} learned flags={flags}, all {wl.T} slides")
    xs = wl.data()
    want = both_paths(xs, wl.W, wl.S, wl.T, wl.R, wl.K, f"{wl.name} learned",
                      sched_kind=R.SCHED_LEARNED, flags=flags)
    assert sum(len(w) for w in want) > 0


def _counters(xs, W, S, T, Rr, K, pipe, frozen=None, **kw):
    dim = xs.shape[1]
    with R.DeviceStream(xs[: W + T * S]) as ds, R.Detector(dim, W, S, Rr, K, **kw) as det:
        if frozen is not None:
            det.set_predictor(frozen, frozen=True)
        det.prime(xs[:W])
        if pipe:
            lists = det.slides_device(ds.at(W), T)
        else:
            lists = [det.slide_device(ds.at(W + i * S)) for i in range(T)]
        st = det.stats()
    return lists, st


@pytest.mark.parametrize("sched", ["plain", "learned"])
def test_pipelined_counters_equal_sequential(sched):
    """Pair tests (and, learned, misses and fallbacks) of a pipelined run equal
    those of the same slides one per call: every slide keeps its own counters,
    and the learned fallback decision of slide t is made by slide t - 2.  The
    learned predictor is frozen (trained weights would depend on timing)."""
    wl = O.C3
    case(f"counters pipelined vs sequential, {sched}")
    xs = wl.data()
    kw = {}
    frozen = None
    if sched == "learned":
        kw["sched_kind"] = R.SCHED_LEARNED
        kw["flags"] = R.FLAG_LEARN_ALWAYS  # C3 alone would run the plain kernels
        # a trained predictor, so that the near-first path is taken
        with R.Detector(3, wl.W, wl.S, wl.R, wl.K, **kw) as det:
            det.prime(xs[:wl.W])
            for t in range(1, 61):
                det.slide(xs[wl.W + (t - 1) * wl.S: wl.W + t * wl.S])
            frozen = det.get_predictor()
    T = 40
    a, sa = _counters(xs, wl.W, wl.S, T, wl.R, wl.K, False, frozen, **kw)
    b, sb = _counters(xs, wl.W, wl.S, T, wl.R, wl.K, True, frozen, **kw)
    assert_same(b, a, "pipelined vs sequential")
    for k in ("pair_tests_total", "candidates_total", "outliers_total", "predictor_misses",
              "fallbacks"):
        assert sa[k] == sb[k], (k, sa[k], sb[k])
    assert sa["pair_tests_total"] > 0


@pytest.mark.parametrize("index", ["lbvh", "stencil", "runs"])
def test_failed_pipelined_run_needs_prime(index, monkeypatch):
    """A pipelined run that fails part-way (a failure injected on the host
    while collecting slide 3) leaves the context unprimed: further slides are
    refused with RTOD_ERR_STATE until rtod_prime, after which the results are
    exact again."""
    case(f"failed pipelined run, {index}")
    W, S, T, Rr, K = 4000, 200, 8, 0.45, 20
    xs = O.gen_walk(77, W + 2 * T * S)
    kind = {"lbvh": R.INDEX_LBVH, "stencil": R.INDEX_CELL_STENCIL,
            "runs": R.INDEX_SLIDE_RUNS}[index]
    monkeypatch.setenv("RTOD_FAIL_AT_SLIDE", "3")
    det = R.Detector(1, W, S, Rr, K, index_kind=kind)
    monkeypatch.delenv("RTOD_FAIL_AT_SLIDE")
    with R.DeviceStream(xs) as ds, det:
        det.prime(xs[:W])
        with pytest.raises(R.RtodError) as e:
            det.slides_device(ds.at(W), T)
        assert "injected" in str(e.value)
        with pytest.raises(R.RtodError) as e2:
            det.slide(xs[W: W + S])
        assert e2.value.status == 3
        with pytest.raises(R.RtodError) as e3:
            det.slides_device(ds.at(W), 2)
        assert e3.value.status == 3
        det.prime(xs[:W])  # slides since prime restart at 0: slide 3 fails again
        got = det.slides_device(ds.at(W), 3)
        assert_same(got, O.run_stream(xs, W, S, 3, Rr, K, "grid"), f"{index} after prime")


@pytest.mark.parametrize("wl,index", [(O.C1, R.INDEX_AUTO), (O.C2, R.INDEX_AUTO),
                                      (O.C2, R.INDEX_CELL_STENCIL), (O.C3, R.INDEX_AUTO)],
                         ids=["C1", "C2", "C2-stencil", "C3"])
def test_host_fed_pipelined_every_slide(wl, index):
    """rtod_slides: the slides from host memory, staged into pinned memory
    just before their ingest, pipelined; all slides of the stream, split into
    runs of uneven length mixed with single calls on one context."""
    case(f"{wl.name} host-fed pipelined, {wl.T} slides, index {index}")
    xs = wl.data()
    want = O.run_stream(xs, wl.W, wl.S, wl.T, wl.R, wl.K, "grid")
    got = []
    with R.Detector(wl.dim, wl.W, wl.S, wl.R, wl.K, index_kind=index) as det:
        det.prime(xs[:wl.W])
        t, runs = 0, [1, 2, 3, 9, 60]
        k = 0
        while t < wl.T:
            n = min(runs[k % len(runs)], wl.T - t)
            k += 1
            rows = xs[wl.W + t * wl.S: wl.W + (t + n) * wl.S]
            if k % 4 == 0:  # a single call in between
                got.append(det.slide(rows[: wl.S]))
                n = 1
            else:
                got.extend(det.slides(rows, n))
            t += n
    assert_same(got, want, f"{wl.name} host-fed pipelined")


def test_host_fed_pipelined_long_run_lbvh():
    """The non-stage pipelined path (LBVH index) from host memory."""
    wl = O.C3
    case("C3 host-fed pipelined, LBVH")
    xs = wl.data()
    T = 40
    with R.Detector(3, wl.W, wl.S, wl.R, wl.K, index_kind=R.INDEX_LBVH) as det:
        det.prime(xs[:wl.W])
        got = det.slides(xs[wl.W: wl.W + T * wl.S], T)
    assert_same(got, O.run_stream(xs, wl.W, wl.S, T, wl.R, wl.K, "grid"), "C3 LBVH host-fed")


@pytest.mark.parametrize("index", ["stencil", "lbvh", "runs"])
def test_time_breakdown_on_every_call(index):
    """The reference's five buckets are filled on every rtod_slide call (no
    flag), from the device stamps of the slide's kernels; with the launch
    remainder they add up to the host wall time."""
    wl = O.C3 if index == "lbvh" else O.C2
    case(f"time breakdown, {index}")
    xs = wl.data()
    kind = {"lbvh": R.INDEX_LBVH, "stencil": R.INDEX_CELL_STENCIL,
            "runs": R.INDEX_SLIDE_RUNS}[index]
    with R.Detector(wl.dim, wl.W, wl.S, wl.R, wl.K, index_kind=kind) as det:
        det.prime(xs[:wl.W])
        for t in range(1, 21):
            det.slide(xs[wl.W + (t - 1) * wl.S: wl.W + t * wl.S])
            tm = det.times
            parts = [tm.copy_h2d, tm.prepare_cell, tm.build_index, tm.detect_outlier, tm.copy_d2h]
            assert all(p > 0 for p in parts), tm
            assert tm.launch >= 0, tm
            assert abs(sum(parts) + tm.launch - tm.total) < 1e-9, tm
            st = det.stage_times()
            assert st["ingest"] > 0 and st["classify"] > 0 and st["detect"] > 0, st
            assert (st["lbvh_build"] > 0) == (index == "lbvh"), st
            assert sum(st.values()) < tm.total, (st, tm)
