"""GPU parity: the HIP path (through the C-ABI) against the CPU oracle.

Bar: bit-exact outlier sets on every slide (ring slots, ascending).  Oracles:
brute force (small windows), the grid oracle (same algorithm, any size) and
the 1-D sorted-window oracle (independent method, full C2 size).
"""
import numpy as np
import pytest

import oracle as O
import rtod_amd as R
from rtod_amd.synth import calibrate_R

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


def test_small_walk_vs_brute():
    case("walk1d W=2000 S=100 K=10")
    W, S, T = 2000, 100, 40
    xs = O.gen_walk(5, W + T * S)
    assert_same(R.run_stream(xs, W, S, T, 0.45, 10), O.run_stream(xs, W, S, T, 0.45, 10, "brute"),
                "walk1d small")


def test_c1_full_stream_vs_grid_and_brute():
    wl = O.C1
    case(wl.name)
    xs = wl.data()
    got = R.run_stream(xs, wl.W, wl.S, wl.T, wl.R, wl.K)
    assert_same(got, O.run_stream(xs, wl.W, wl.S, wl.T, wl.R, wl.K, "grid"), "C1 vs grid")
    assert_same(got[:20], O.run_stream(xs, wl.W, wl.S, 20, wl.R, wl.K, "brute"), "C1 vs brute")


def test_c2_vs_sorted1d_and_grid():
    wl = O.C2
    case(wl.name)
    xs = wl.data()
    T = 40
    got = R.run_stream(xs, wl.W, wl.S, T, wl.R, wl.K)
    assert_same(got, O.run_stream(xs, wl.W, wl.S, T, wl.R, wl.K, "grid"), "C2 vs grid")
    assert_same(got[:10], O.run_stream(xs, wl.W, wl.S, 10, wl.R, wl.K, "sorted1d"),
                "C2 vs sorted1d")


def test_c3_vs_grid_and_brute():
    wl = O.C3
    case(wl.name)
    xs = wl.data()
    T = 120
    got = R.run_stream(xs, wl.W, wl.S, T, wl.R, wl.K)
    assert_same(got, O.run_stream(xs, wl.W, wl.S, T, wl.R, wl.K, "grid"), "C3 vs grid")
    assert_same(got[:5], O.run_stream(xs, wl.W, wl.S, 5, wl.R, wl.K, "brute"), "C3 vs brute")


@pytest.mark.parametrize("kind", ["walk", "gmm"])
def test_2d_vs_brute(kind):
    case(f"2-D {kind}")
    W, S, T = 3000, 150, 20
    xs = O.gen_walk(21, W + T * S, 2) if kind == "walk" else O.gen_gmm(22, W + T * S, 2)
    Rr = 0.3 if kind == "walk" else 1.5
    assert_same(R.run_stream(xs, W, S, T, Rr, 20), O.run_stream(xs, W, S, T, Rr, 20, "brute"),
                f"2-D {kind}")


@pytest.mark.parametrize("dim,W,S", [(1, 2050, 1025), (2, 12291, 4097), (3, 16384, 8192),
                                     (1, 24576, 8192)])
def test_multi_tile_batches_vs_grid(dim, W, S):
    # batches of several 1024-point sort tiles, merged in one pass (two runs
    # with a one-point tail, five runs, eight full runs), priming included
    case(f"multi-tile dim={dim} W={W} S={S}")
    T = 4
    xs = O.gen_walk(61 + dim, W + T * S, dim) if dim < 3 else O.gen_gmm(64, W + T * S, 3)
    Rr = {1: 0.45, 2: 0.4, 3: 1.9}[dim]
    got = R.run_stream(xs, W, S, T, Rr, 30)
    assert_same(got, O.run_stream(xs, W, S, T, Rr, 30, "grid"), f"multi-tile dim={dim} S={S}")


@pytest.mark.parametrize("W,S,K", [(1000, 1000, 5), (999, 1, 3), (4096, 64, 4096),
                                   (5000, 2500, 1), (40000, 20000, 50)])
def test_shapes_vs_oracle(W, S, K):
    case(f"shape W={W} S={S} K={K}")
    T = 3 if S > 1 else 50
    xs = O.gen_walk(31, W + T * S)
    want = O.run_stream(xs, W, S, T, 0.45, K, "sorted1d")
    assert_same(R.run_stream(xs, W, S, T, 0.45, K), want, f"W={W} S={S} K={K}")


def test_lattice_closed_form():
    """x_i = i*h, h = 0.25, R = 0.5 exactly: interior points have exactly
    2*R/h + 1 = 5 neighbours (<= includes distance R); K=4 flags only the
    window's edge-free lattice ends; K=5 flags every point."""
    case("lattice")
    h, W, S, T = 0.25, 256, 64, 6
    xs = (np.arange(W + T * S, dtype=np.float64) * h).astype(np.float32).reshape(-1, 1)
    for K in (4, 5):
        got = R.run_stream(xs, W, S, T, 0.5, K)
        want = O.run_stream(xs, W, S, T, 0.5, K, "brute")
        assert_same(got, want, f"lattice K={K}")
        if K == 5:
            assert all(len(g) == W for g in got)
        else:  # the two lowest and two highest values of each window
            assert all(len(g) == 4 for g in got)


def test_duplicates_and_far_points():
    case("duplicates")
    W, S, T, K = 1200, 100, 12, 30
    rng = np.random.default_rng(7)
    xs = rng.integers(0, 40, size=(W + T * S, 3)).astype(np.float32) * 0.25
    xs[::97] += 1000.0  # isolated far points: always outliers
    got = R.run_stream(xs, W, S, T, 0.3, K)
    assert_same(got, O.run_stream(xs, W, S, T, 0.3, K, "brute"), "duplicates")


def test_nonfinite_points():
    """NaN/inf points have no neighbours, themselves included (d2 is NaN or
    inf), so they are outliers and count for nobody -- as in brute force."""
    case("nonfinite")
    W, S, T, K = 800, 100, 6, 5
    xs = O.gen_walk(41, W + T * S, 2)
    xs[17] = [np.nan, 1.0]
    xs[300] = [np.inf, 2.0]
    xs[905] = [-np.inf, -np.inf]
    xs[1200] = [np.nan, np.nan]
    got = R.run_stream(xs, W, S, T, 0.45, K)
    assert_same(got, O.run_stream(xs, W, S, T, 0.45, K, "brute"), "nonfinite")


def test_huge_coordinates_clamped_cells():
    """Coordinates beyond the cell-key range land in clamped cells, which are
    never fast-accepted; results stay exact."""
    case("clamped")
    W, S, T, K = 600, 60, 5, 3
    rng = np.random.default_rng(9)
    xs = np.concatenate([rng.normal(0, 1, (W + T * S, 1)),
                         rng.normal(0, 1, (W + T * S, 1))], 1).astype(np.float32)
    xs[::3, 0] += np.float32(3e9)
    got = R.run_stream(xs, W, S, T, 0.5, K)
    assert_same(got, O.run_stream(xs, W, S, T, 0.5, K, "brute"), "clamped")


def test_wide_keys_two_stage_sort():
    """3-D points at both clamp extremes of every axis make the batch's
    relative cell keys 63 bits wide, which takes the two-stage sort."""
    case("wide keys")
    W, S, T, K = 1500, 500, 6, 4
    rng = np.random.default_rng(10)
    xs = rng.normal(0, 1, (W + T * S, 3)).astype(np.float32)
    sel = rng.integers(0, 3, (W + T * S, 3))
    xs[sel == 1] += np.float32(1e7)
    xs[sel == 2] -= np.float32(1e7)
    got = R.run_stream(xs, W, S, T, 1.0, K)
    assert_same(got, O.run_stream(xs, W, S, T, 1.0, K, "brute"), "wide keys")


def test_phase_api_equals_slide():
    case("phase api")
    W, S, T, K = 5000, 250, 10, 20
    xs = O.gen_gmm(51, W + T * S, 3)
    want = R.run_stream(xs, W, S, T, 1.9, K)
    with R.Detector(3, W, S, 1.9, K) as det:
        det.prime(xs[:W])
        got = []
        for t in range(1, T + 1):
            det.copy_new_points(xs[W + (t - 1) * S: W + t * S])
            det.prepare_cell()
            det.build_index()
            det.detect_outlier()
            got.append(det.copy_outlier())
        with pytest.raises(R.RtodError):
            det.detect_outlier()  # out of order
    assert_same(got, want, "phase api")


@pytest.mark.parametrize("index_kind", [R.INDEX_CELL_STENCIL, R.INDEX_SLIDE_RUNS])
@pytest.mark.parametrize("flags", [0, R.FLAG_PHASE_TIMES])
def test_launch_modes_agree(flags, index_kind):
    """With RTOD_FLAG_PHASE_TIMES the slide's kernels are timed with HIP
    events: the cell stencil's detect, or the slide-run index's one launch
    per slide (a run of S <= 8192 points is sorted and detected by one
    workgroup)."""
    case(f"flags={flags} index={index_kind}")
    wl = O.C1
    xs = wl.data()
    T = 30
    want = O.run_stream(xs, wl.W, wl.S, T, wl.R, wl.K, "grid")
    with R.Detector(1, wl.W, wl.S, wl.R, wl.K, flags=flags, index_kind=index_kind) as det:
        det.prime(xs[:wl.W])
        got = [det.slide(xs[wl.W + (t - 1) * wl.S: wl.W + t * wl.S]) for t in range(1, T + 1)]
        assert det.stats()["index_kind"] == index_kind
        if flags & R.FLAG_PHASE_TIMES:
            kt = det.kernel_times()
            k = "detect" if index_kind == R.INDEX_CELL_STENCIL else "run_slide"
            assert kt[k][1] == T and kt[k][0] > 0
            assert det.times.total > 0
    assert_same(got, want, f"flags={flags}")


def test_pipelined_phase_times_count_full_launches():
    """With RTOD_FLAG_PHASE_TIMES a pipelined run on the slide-run index
    times each launch that detects a batch and sorts the next as run_stage
    and its last launch, which only detects, as run_detect; the lists are
    the oracle's."""
    case("pipelined phase times")
    wl = O.C1
    xs = wl.data()
    n = R.RUN_FIRST_BATCH + 2 * R.RUN_BATCH + 5  # batches 32, 64, 64, 5
    want = O.run_stream(xs, wl.W, wl.S, n, wl.R, wl.K, "grid")
    with R.DeviceStream(xs[: wl.W + n * wl.S]) as ds, \
            R.Detector(1, wl.W, wl.S, wl.R, wl.K, flags=R.FLAG_PHASE_TIMES,
                       index_kind=R.INDEX_SLIDE_RUNS) as det:
        det.prime(xs[:wl.W])
        got = det.slides_device(ds.at(wl.W), n)
        kt = det.kernel_times()
    assert kt["run_stage"][1] == 3 and kt["run_stage"][0] > 0
    assert kt["run_detect"][1] == 1
    assert_same(got, want, "pipelined phase times")


@pytest.mark.parametrize("dim,sched", [(1, R.SCHED_PLAIN), (3, R.SCHED_PLAIN),
                                       (3, R.SCHED_LEARNED)])
def test_pinned_slides_read_in_place(dim, sched):
    """Slides in a buffer from rtod_host_alloc are read by the kernels in
    place, pipelined and one per call, on the slide-run index (1-D) and the
    cell stencil (3-D, plain and learned schedules): the lists are the
    oracle's."""
    case(f"pinned dim={dim} sched={sched}")
    if dim == 1:
        wl = O.C1
        xs, W, S, Rr, K, n = wl.data(), wl.W, wl.S, wl.R, wl.K, 70
    else:
        W, S, Rr, K, n = 3000, 150, 1.9, 20, 24
        xs = O.gen_gmm(81, W + n * S, 3)
    want = O.run_stream(xs, W, S, n, Rr, K, "grid")
    buf = R.pinned_empty(n * S, dim)
    buf[:] = xs[W: W + n * S]
    with R.Detector(dim, W, S, Rr, K, sched_kind=sched) as det:
        det.prime(xs[:W])
        got = det.slides(buf, n)
    assert_same(got, want, f"pinned pipelined dim={dim}")
    with R.Detector(dim, W, S, Rr, K, sched_kind=sched) as det:
        det.prime(xs[:W])
        got = [det.slide(buf[t * S: (t + 1) * S]) for t in range(n)]
    assert_same(got, want, f"pinned one per call dim={dim}")


def test_slide_device_equals_slide():
    case("slide_device")
    W, S, T, K = 6000, 300, 15, 20
    xs = O.gen_gmm(71, W + T * S, 3)
    want = R.run_stream(xs, W, S, T, 1.9, K)
    with R.DeviceStream(xs) as ds, R.Detector(3, W, S, 1.9, K) as det:
        det.prime(xs[:W])
        got = [det.slide_device(ds.at(W + (t - 1) * S)) for t in range(1, T + 1)]
        with pytest.raises(R.RtodError) as e:  # a host pointer is refused
            det.slide_device(xs.ctypes.data)
        assert e.value.status == 1
    assert_same(got, want, "slide_device")


PIPE_CASES = [
    ("C2", O.C2, 40, {}),
    ("C3", O.C3, 60, {}),
    ("C1 learned", O.C1, 60, dict(sched_kind=R.SCHED_LEARNED)),
    ("C3 learned kernels", O.C3, 40,
     dict(sched_kind=R.SCHED_LEARNED, flags=R.FLAG_LEARN_ALWAYS)),
    ("C3 lbvh", O.C3, 12, dict(index_kind=R.INDEX_LBVH)),
]


@pytest.mark.parametrize("name,wl,T,kw", PIPE_CASES, ids=[c[0] for c in PIPE_CASES])
def test_pipelined_slides_vs_grid(name, wl, T, kw):
    """rtod_slides_device (stage launches on one stream: one per slide, two
    with the learned schedule; the slides' sorts, merges and lists riding in
    neighbouring slides' launches) against
    the grid oracle, mixed with single slides on the same context: 3 single,
    T-8 pipelined, 2 single, 3 pipelined."""
    case(f"pipelined {name}")
    xs = wl.data(wl.W + T * wl.S)
    want = O.run_stream(xs, wl.W, wl.S, T, wl.R, wl.K, "grid")
    W, S = wl.W, wl.S
    got = []
    with R.DeviceStream(xs) as ds, R.Detector(wl.dim, W, S, wl.R, wl.K, **kw) as det:
        det.prime(xs[:W])
        t = 0

        def single(n):
            nonlocal t
            for _ in range(n):
                got.append(det.slide_device(ds.at(W + t * S)))
                t += 1

        def piped(n):
            nonlocal t
            got.extend(det.slides_device(ds.at(W + t * S), n))
            t += n

        single(3)
        piped(T - 8)
        single(2)
        piped(3)
        assert det.stats()["slides"] == T
    assert_same(got, want, f"pipelined {name}")


@pytest.mark.parametrize("dim,W,S,K,kind", [
    (1, 20000, 2500, 20, "walk"),   # three sort tiles: run merge two slides ahead
    (3, 8000, 400, 20, "gmm"),      # one sort tile: sorted straight into the ring
    (2, 6000, 1500, 20, "walk"),    # two tiles
    (1, 36864, 12288, 20, "walk"),  # six 2048-point tiles
    (3, 32768, 16384, 20, "gmm"),   # eight 2048-point tiles: the run merge's cap
])
def test_pipelined_short_runs(dim, W, S, K, kind):
    """Pipelined calls of 1, 2, 3, 1, 4 and 5 slides in a row (every length
    of prologue and tail: the sorts and merges run up to three slides ahead
    of the detect, the list one slide behind) equal the grid oracle."""
    case(f"pipelined short runs dim={dim} W={W} S={S}")
    runs = [1, 2, 3, 1, 4, 5]
    T = sum(runs)
    xs = O.gen_walk(90 + dim, W + T * S, dim) if kind == "walk" else \
        O.gen_gmm(90 + dim, W + T * S, dim)
    Rr = calibrate_R(xs[:W], K, seed=dim)
    want = O.run_stream(xs, W, S, T, Rr, K, "grid")
    got = []
    with R.DeviceStream(xs) as ds, R.Detector(dim, W, S, Rr, K) as det:
        det.prime(xs[:W])
        t = 0
        for n in runs:
            got.extend(det.slides_device(ds.at(W + t * S), n))
            t += n
        assert det.stats()["slides"] == T
    assert_same(got, want, f"short runs dim={dim}")


def test_pipelined_slides_argument_checks():
    case("pipelined args")
    W, S = 2000, 200
    xs = O.gen_walk(81, W + 4 * S)
    with R.DeviceStream(xs) as ds, R.Detector(1, W, S, 0.45, 10) as det:
        offs = np.zeros(3, np.int32)
        slots = np.zeros(2 * W, np.int32)
        L = R.lib()
        # before prime
        assert L.rtod_slides_device(det._h, ds.at(W), 2, offs, slots, slots.size) == 3
        det.prime(xs[:W])
        assert L.rtod_slides_device(det._h, ds.at(W), 2, offs, slots, slots.size - 1) == 1
        assert L.rtod_slides_device(det._h, ds.at(W), 0, offs, slots, slots.size) == 1
        assert L.rtod_slides_device(det._h, xs.ctypes.data, 2, offs, slots, slots.size) == 1
        # nothing ran: the stream continues from slide 1
        assert det.stats()["slides"] == 0
        got = det.slides_device(ds.at(W), 2)
    assert_same(got, O.run_stream(xs, W, S, 2, 0.45, 10, "grid"), "pipelined after errors")


def test_classify_wait_timeout_fails_slides(monkeypatch):
    """A pipelined slide's detect waits for the classify of its own launch
    with a bounded spin.  A context created with RTOD_CLASSIFY_WAIT_SPINS=0
    gives up at once: the pipelined call fails with error flag 8 (every
    kernel still exits), and after priming again the context's one-slide
    calls equal the grid oracle."""
    case("classify wait bound 0")
    wl, T = O.C1, 6
    W, S = wl.W, wl.S
    xs = wl.data(W + T * S)
    want = O.run_stream(xs, W, S, T, wl.R, wl.K, "grid")
    monkeypatch.setenv("RTOD_CLASSIFY_WAIT_SPINS", "0")
    with R.DeviceStream(xs) as ds, R.Detector(1, W, S, wl.R, wl.K,
                                              index_kind=R.INDEX_CELL_STENCIL) as det:
        monkeypatch.delenv("RTOD_CLASSIFY_WAIT_SPINS")
        det.prime(xs[:W])
        with pytest.raises(R.RtodError, match="code 8"):
            det.slides_device(ds.at(W), 4)
        det.prime(xs[:W])
        got = [det.slide_device(ds.at(W + t * S)) for t in range(T)]
    assert_same(got, want, "one-slide calls after a failed wait")


def test_reprime_restarts_stream():
    case("reprime")
    W, S, T, K = 2000, 200, 5, 10
    xs = O.gen_walk(61, W + T * S)
    ys = O.gen_walk(62, W + T * S)
    with R.Detector(1, W, S, 0.45, K) as det:
        det.prime(xs[:W])
        for t in range(1, 3):
            det.slide(xs[W + (t - 1) * S: W + t * S])
        det.prime(ys[:W])
        got = [det.slide(ys[W + (t - 1) * S: W + t * S]) for t in range(1, T + 1)]
    assert_same(got, O.run_stream(ys, W, S, T, 0.45, K, "brute"), "reprime")


def test_stats_counters():
    case("stats")
    wl = O.C1
    xs = wl.data()
    g = O.Grid(1, wl.W, wl.S, wl.R, wl.K)
    g.prime(xs[:wl.W])
    with R.Detector(1, wl.W, wl.S, wl.R, wl.K) as det:
        det.prime(xs[:wl.W])
        for t in range(1, 6):
            new = xs[wl.W + (t - 1) * wl.S: wl.W + t * wl.S]
            out = det.slide(new)
            g.slide(new)
        st = det.stats()
    assert st["slides"] == 5
    assert st["outliers_last"] == len(out)
    # same cells (same key function) -> same candidate set size
    assert st["candidates_last"] == g.stats()["candidates"]
    assert st["device_bytes"] > 0 and st["host_pinned_bytes"] > 0


# ---------------------------------------------------------------- LBVH index
LBVH_CASES = [
    ("walk1d W=2000", lambda: O.gen_walk(5, 2000 + 30 * 100), 2000, 100, 30, 0.45, 10),
    ("gmm3d W=5000", lambda: O.gen_gmm(13, 5000 + 20 * 250, 3), 5000, 250, 20, 1.9, 50),
    ("gmm2d W=3000", lambda: O.gen_gmm(22, 3000 + 20 * 150, 2), 3000, 150, 20, 1.5, 20),
    ("walk3d W=2400", lambda: O.gen_walk(7, 2400 + 15 * 120, 3), 2400, 120, 15, 0.3, 20),
]


@pytest.mark.parametrize("name,gen,W,S,T,Rr,K", LBVH_CASES, ids=[c[0] for c in LBVH_CASES])
def test_lbvh_vs_brute(name, gen, W, S, T, Rr, K):
    case(f"lbvh {name}")
    xs = gen()
    got = R.run_stream(xs, W, S, T, Rr, K, index_kind=R.INDEX_LBVH)
    assert_same(got, O.run_stream(xs, W, S, T, Rr, K, "brute"), f"lbvh {name}")


def test_lbvh_c3_vs_grid():
    wl = O.C3
    case("lbvh " + wl.name)
    xs = wl.data()
    T = 30
    got = R.run_stream(xs, wl.W, wl.S, T, wl.R, wl.K, index_kind=R.INDEX_LBVH)
    assert_same(got, O.run_stream(xs, wl.W, wl.S, T, wl.R, wl.K, "grid"), "lbvh C3")


def test_lbvh_c2_multi_tile_vs_grid():
    """C2 with R small enough that candidates exceed one 4096-entry sort tile
    (merge passes of the LBVH sort run)."""
    wl = O.C2
    case("lbvh C2 small R")
    xs = wl.data()
    T, Rr, K = 4, 0.02, 200
    got = R.run_stream(xs, wl.W, wl.S, T, Rr, K, index_kind=R.INDEX_LBVH)
    with R.Detector(1, wl.W, wl.S, Rr, K, index_kind=R.INDEX_LBVH) as det:
        det.prime(xs[:wl.W])
        det.slide(xs[wl.W: wl.W + wl.S])
        assert det.stats()["candidates_last"] > 4096
    assert_same(got, O.run_stream(xs, wl.W, wl.S, T, Rr, K, "grid"), "lbvh C2 small R")


@pytest.mark.parametrize("index_kind", [R.INDEX_CELL_STENCIL, R.INDEX_LBVH])
def test_no_candidates_and_edge_data(index_kind):
    case(f"edge index={index_kind}")
    W, S, T = 1000, 100, 4
    same = np.zeros((W + T * S, 2), np.float32)  # one dense cell: no candidates
    got = R.run_stream(same, W, S, T, 0.5, 10, index_kind=index_kind)
    assert all(len(g) == 0 for g in got)
    xs = O.gen_walk(41, W + T * S, 2)
    xs[17] = [np.nan, 1.0]
    xs[300] = [np.inf, 2.0]
    xs[905] = [-np.inf, -np.inf]
    got = R.run_stream(xs, W, S, T, 0.45, 5, index_kind=index_kind)
    assert_same(got, O.run_stream(xs, W, S, T, 0.45, 5, "brute"), "edge data")
