"""GPU parity of the 1-D slide-run index (RTOD_INDEX_SLIDE_RUNS, the default
for 1-D streams) through the C-ABI, against the oracle's brute force or grid.

Each case drives one of the index's modes or limits: runs whose cell range
exceeds the direct-mode bucket table (searched by cell), windows whose cell
range exceeds the LDS histogram (cells counted by search), more cell-table
entries and more candidates than one round holds, every point an outlier,
one run per window, two runs per slide, lattices with R on a lattice
distance (the FP32 test decides at the boundary buckets), non-finite and
clamped values, pipelined runs of every length around the launch batch, and
the index's eligibility rules.  Bar: bit-exact slot sets on every slide,
one per call and pipelined.
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


def check(xs, W, S, T, Rr, K, name, method="brute"):
    """One call per slide and pipelined (host-fed) against the oracle; the
    context must be on the slide-run index."""
    want = O.run_stream(xs, W, S, T, Rr, K, method)
    with R.Detector(1, W, S, Rr, K) as det:
        assert det.stats()["index_kind"] == R.INDEX_SLIDE_RUNS
        det.prime(xs[:W])
        got = [det.slide(xs[W + t * S: W + (t + 1) * S]) for t in range(T)]
        assert_same(got, want, f"{name} one per call")
        det.prime(xs[:W])
        assert_same(det.slides(xs[W:W + T * S], T), want, f"{name} pipelined")
    return want


def test_search_mode_runs():
    """Each slide spans ~20k cells (more than the 4096 buckets of a direct
    run): every run is sorted by cell and searched."""
    case("search-mode runs")
    rng = np.random.default_rng(11)
    W, S, T = 4000, 400, 12
    xs = rng.uniform(0, 1e4, (W + T * S, 1)).astype(np.float32)
    want = check(xs, W, S, T, 0.45, 3, "search mode")
    assert sum(len(w) for w in want) > 0


def test_window_cell_range_beyond_histogram():
    """Each slide is compact (a direct run) but the slides sit 1000 units
    apart, so the window spans more cells than the LDS histogram: the
    window's cell counts come from searching every run's cell table."""
    case("window beyond the LDS histogram")
    rng = np.random.default_rng(12)
    W, S, T = 3000, 300, 12
    n = W + T * S
    xs = (np.repeat(np.arange(n // S) * 1000.0, S) + rng.normal(0, 2, n)).astype(np.float32)
    xs[::97] += 0.3  # a few points shifted onto their neighbours' cells
    check(xs[:, None], W, S, T, 0.45, 5, "wide window")


def test_many_entries_and_candidates():
    """About 3000 cells per run and every cell sparse: more cell-table
    entries than one LDS round (2048) and ~20k candidates, counted in
    rounds of 2048."""
    case("many entries and candidates")
    rng = np.random.default_rng(13)
    W, S, T = 20000, 2000, 6
    xs = rng.uniform(0, 1500, (W + T * S, 1)).astype(np.float32)
    check(xs, W, S, T, 0.45, 50, "many candidates", method="grid")


def test_every_point_an_outlier_and_one_run_window():
    case("K >= W, and S = W")
    rng = np.random.default_rng(14)
    W, S, T = 1000, 100, 5
    xs = rng.normal(0, 1, (W + T * S, 1)).astype(np.float32)
    want = check(xs, W, S, T, 0.3, W + 5, "K >= W")
    assert all(len(w) == W for w in want)
    W, S, T = 512, 512, 6
    xs = rng.normal(0, 1, (W + T * S, 1)).astype(np.float32)
    check(xs, W, S, T, 0.2, 10, "S = W")


def test_outlier_counts_around_the_ranked_list():
    """Windows whose outlier count steps through 256: a slide with up to 256
    outliers lists them by rank, one with more from the bitmap.  Each slide
    holds a dense cluster (inliers) and isolated points 0.5 apart (outliers,
    each alone within R = 0.45), their number rising every other slide so
    that the window's count goes from 250 to 340; the window stays within
    the one-slide workgroup's cell histogram."""
    case("outlier counts around 256")
    rng = np.random.default_rng(22)
    W, S, T = 2000, 200, 24
    m = W // S
    slides, nxt = [], 0
    for t in range(m + T):
        k = 25 + max(0, t - m) // 2  # isolated points in slide t
        iso = 2.0 + 0.5 * np.arange(nxt, nxt + k)
        nxt += k
        dense = rng.uniform(0.0, 0.1, S - k)
        pts = np.concatenate([iso, dense])
        slides.append(rng.permutation(pts))
    xs = np.concatenate(slides).astype(np.float32)[:, None]
    want = check(xs, W, S, T, 0.45, 5, "ranked list")
    counts = [len(w) for w in want]
    assert min(counts) <= 256 < max(counts), counts


@pytest.mark.parametrize("S", [1, 2, 3, 5, 7, 13])
def test_tiny_and_ragged_slides(S):
    """Slides of fewer than four points (no full 16-byte group: the group
    loads are discarded) and of lengths that leave a tail of one to three
    points after the groups."""
    case(f"slides of {S} points")
    rng = np.random.default_rng(20 + S)
    W, T = 64 * S, 12
    xs = np.cumsum(rng.normal(0, 0.2, W + T * S)).astype(np.float32)[:, None]
    check(xs, W, S, T, 0.45, 4, f"S={S}")


def test_two_runs_per_slide():
    """S = 16384 (two runs of 8192 per slide): the one-per-call path is a
    sort launch and a detect launch."""
    case("two runs per slide")
    wl = O.C2
    W, S, T = 65536, 16384, 5
    xs = wl.data(W + T * S)
    check(xs, W, S, T, 0.45, 50, "S=16384", method="grid")


@pytest.mark.parametrize("step,Rr,K", [(0.1, 0.3, 24), (0.25, 0.5, 16), (1.0 / 3.0, 1.0, 24)])
def test_lattice_boundaries(step, Rr, K):
    """Points on a lattice with R at (or within rounding of) a lattice
    distance, each value four times in the window: a point has 4 (2j + 1)
    neighbours with j the lattice steps within R, so with K between the
    counts the outlier status of every point turns on whether the FP32 test
    passes at exactly R, decided inside the boundary buckets (no cell is
    dense: it holds fewer than K + 1 points)."""
    case(f"lattice step {step} R {Rr} K {K}")
    W, S, T = 2400, 240, 8
    n = W + T * S
    base = (np.arange(n) % 600) * step
    xs = np.asarray(base, np.float32)[:, None]
    want = check(xs, W, S, T, Rr, K, f"lattice {step}")
    assert 0 < sum(len(w) for w in want) < T * W
    rng = np.random.default_rng(15)
    xs2 = np.asarray(rng.permutation(base), np.float32)[:, None]
    check(xs2, W, S, T, Rr, K, f"lattice {step} shuffled")


def test_nonfinite_clamped_and_duplicates():
    case("NaN, inf, clamped, duplicates")
    rng = np.random.default_rng(16)
    W, S, T = 3000, 300, 10
    xs = rng.normal(0, 1, (W + T * S, 1)).astype(np.float32)
    n = len(xs)
    xs[rng.integers(0, n, 25)] = np.nan
    xs[rng.integers(0, n, 10)] = np.inf
    xs[rng.integers(0, n, 10)] = -np.inf
    xs[rng.integers(0, n, 10)] = 3e38
    xs[rng.integers(0, n, 10)] = -1e30
    xs[rng.integers(0, n, 60)] = 5.0
    check(xs, W, S, T, 0.3, 10, "edge data")


@pytest.mark.parametrize("n", [1, 31, 32, 33, 64, 95, 96, 97, 159, 160, 161, 225])
def test_pipelined_run_lengths(n):
    """Runs of n slides around the batches (the first 32 slides, then 32 a
    batch from host memory and 128 from HBM; a launch detects one batch and
    sorts the next), host-fed and from HBM, then single calls on the same
    context."""
    case(f"pipelined run of {n}")
    wl = O.C1
    T = n + 3
    xs = wl.data(wl.W + T * wl.S)
    want = O.run_stream(xs, wl.W, wl.S, T, wl.R, wl.K, "grid")
    with R.DeviceStream(xs) as ds, R.Detector(1, wl.W, wl.S, wl.R, wl.K) as det:
        det.prime(xs[:wl.W])
        got = det.slides(xs[wl.W:wl.W + n * wl.S], n)
        got += [det.slide(xs[wl.W + t * wl.S: wl.W + (t + 1) * wl.S]) for t in range(n, T)]
        assert_same(got, want, f"host-fed run of {n}")
        det.prime(xs[:wl.W])
        got = det.slides_device(ds.at(wl.W), n)
        got += [det.slide_device(ds.at(wl.W + t * wl.S)) for t in range(n, T)]
        assert_same(got, want, f"HBM run of {n}")


def test_eligibility():
    """AUTO picks the slide-run index only where it applies; requested where
    it does not, rtod_create fails with RTOD_ERR_UNSUPPORTED."""
    case("eligibility")
    with R.Detector(1, 10000, 500, 0.45, 50) as det:
        assert det.stats()["index_kind"] == R.INDEX_SLIDE_RUNS
    with R.Detector(1, 1 << 19, 1 << 13, 0.45, 50) as det:  # W > 2^18: the cell stencil
        assert det.stats()["index_kind"] == R.INDEX_CELL_STENCIL
    with R.Detector(1, 10000, 20, 0.45, 50) as det:  # 500 runs per window
        assert det.stats()["index_kind"] == R.INDEX_CELL_STENCIL
    with R.Detector(1, 10000, 500, 0.45, 50, sched_kind=R.SCHED_LEARNED,
                    flags=R.FLAG_LEARN_ALWAYS) as det:
        assert det.stats()["index_kind"] == R.INDEX_CELL_STENCIL
    for kw in (dict(dim=3), dict(window=1 << 19, slide=1 << 13),
               dict(ablation=R.ABL_NO_FILTER)):
        args = dict(dim=1, window=10000, slide=500)
        args.update(kw)
        with pytest.raises(R.RtodError) as e:
            R.Detector(args["dim"], args["window"], args["slide"], 0.45, 50,
                       index_kind=R.INDEX_SLIDE_RUNS, ablation=kw.get("ablation", 0))
        assert e.value.status == 2, kw


def test_learned_context_runs_plain():
    """A 1-D learned context (cost rule) runs on the slide-run index: every
    slide counts as run on the plain kernels and as a fallback, the
    predictor API still works, and the outliers equal the oracle's."""
    case("learned 1-D context")
    wl = O.C1
    T = 20
    xs = wl.data(wl.W + T * wl.S)
    want = O.run_stream(xs, wl.W, wl.S, T, wl.R, wl.K, "grid")
    with R.Detector(1, wl.W, wl.S, wl.R, wl.K, sched_kind=R.SCHED_LEARNED) as det:
        p = det.get_predictor()
        det.set_predictor(p, frozen=True)
        assert np.array_equal(det.get_predictor(), p)
        det.prime(xs[:wl.W])
        got = [det.slide(xs[wl.W + t * wl.S: wl.W + (t + 1) * wl.S]) for t in range(T)]
        st = det.stats()
    assert_same(got, want, "learned 1-D")
    assert st["index_kind"] == R.INDEX_SLIDE_RUNS
    assert st["plain_kernel_slides"] == T and st["fallbacks"] == T
    assert st["train_steps"] == 0


def test_phase_api_and_stage_times():
    """The phase calls on the slide-run index (prepare_cell sorts the run,
    detect_outlier detects) equal rtod_slide; the stage times of a slide
    run one per call cover the sort, the window histogram, the counts and
    the list."""
    case("phase API, stage times")
    wl = O.C1
    T = 12
    xs = wl.data(wl.W + T * wl.S)
    want = O.run_stream(xs, wl.W, wl.S, T, wl.R, wl.K, "grid")
    with R.Detector(1, wl.W, wl.S, wl.R, wl.K) as det:
        det.prime(xs[:wl.W])
        got = []
        for t in range(T):
            new = xs[wl.W + t * wl.S: wl.W + (t + 1) * wl.S]
            if t % 2:
                got.append(det.slide(new))
                st = det.stage_times()
                assert st["ingest"] > 0 and st["classify"] > 0 and st["detect"] > 0, st
                assert st["epilogue"] > 0 and st["lbvh_build"] == 0, st
            else:
                det.copy_new_points(new)
                det.prepare_cell()
                det.build_index()
                det.detect_outlier()
                got.append(det.copy_outlier())
    assert_same(got, want, "phase API")
