"""GPU parity of the ablation modes, the reference's OPTIMIZATION 0/1,
UPDATE_GAS_TYPE 0 and --launch_ray_num (optixScan.h:18-24, optixScan.cu:80-156,
optixScan.cpp:200-206, 335-376, 738-745), through the C-ABI.

The outlier set of a window does not depend on the mode, only the work done:
every mode is checked against the grid oracle on every slide it runs.  With
launch_ray_num L the reference casts ray i from window slot i mod W and
lists the rays whose count is <= K, so the expected list is the oracle's
slots below L (L <= W), or those slots repeated every W below L (L > W).
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


def rays(slots, W, L):
    """The reference's ray-index list for launch_ray_num L."""
    if L == 0 or L == W:
        return slots
    if L < W:
        return slots[slots < L]
    out = np.concatenate([slots + b for b in range(0, L, W)])
    return np.sort(out[out < L]).astype(np.int32)


STREAMS = {
    "walk1d": lambda: (O.C1.data(), O.C1.W, O.C1.S, O.C1.R, O.C1.K),
    "gmm3d": lambda: (O.C3.data(), O.C3.W, O.C3.S, O.C3.R, O.C3.K),
    "gmm2d": lambda: (O.gen_gmm(91, 6000 + 60 * 300, 2), 6000, 300, 1.5, 20),
}

MODES = [
    # (index, ablation, refit)
    ("stencil", R.ABL_FILTER_ONLY, False),
    ("stencil", R.ABL_NO_FILTER, False),
    ("lbvh", R.ABL_FILTER_ONLY, False),
    ("lbvh", R.ABL_FILTER_ONLY, True),
    ("lbvh", R.ABL_NO_FILTER, False),
    ("lbvh", R.ABL_NO_FILTER, True),
]


def _kw(index, abl, refit, L=0):
    return dict(index_kind=R.INDEX_LBVH if index == "lbvh" else R.INDEX_CELL_STENCIL,
                ablation=abl, refit=refit, launch_ray_num=L)


@pytest.mark.parametrize("stream", sorted(STREAMS))
@pytest.mark.parametrize("index,abl,refit", MODES)
def test_ablation_vs_grid(stream, index, abl, refit):
    xs, W, S, Rr, K = STREAMS[stream]()
    T = 60 if stream != "walk1d" else 120
    case(f"{stream} {index} ablation={abl} refit={refit}, {T} slides")
    want = O.run_stream(xs, W, S, T, Rr, K, "grid")
    kw = _kw(index, abl, refit)
    with R.Detector(xs.shape[1], W, S, Rr, K, **kw) as det:
        det.prime(xs[:W])
        got = [det.slide(xs[W + (t - 1) * S: W + t * S]) for t in range(1, T // 2 + 1)]
        # the rest pipelined on the same context (from host memory)
        got += det.slides(xs[W + (T // 2) * S: W + T * S], T - T // 2)
        st = det.stats()
    assert_same(got, want, f"{stream} {index} abl={abl} refit={refit}")
    if abl == R.ABL_NO_FILTER:  # every window point is a candidate
        assert st["candidates_last"] == W, st


@pytest.mark.parametrize("index", ["stencil", "lbvh"])
@pytest.mark.parametrize("frac", [0.3, 1.0, 2.5])
def test_launch_ray_num(index, frac):
    """--launch_ray_num L with no grid filter: rays from slots i mod W."""
    xs, W, S, Rr, K = STREAMS["walk1d"]()
    L = int(frac * W) + (7 if frac != 1.0 else 0)
    T = 30
    case(f"launch_ray_num={L} (W={W}) {index}")
    want = [rays(w, W, L) for w in O.run_stream(xs, W, S, T, Rr, K, "grid")]
    kw = _kw(index, R.ABL_NO_FILTER, index == "lbvh", L)
    with R.DeviceStream(xs[: W + T * S]) as ds, R.Detector(1, W, S, Rr, K, **kw) as det:
        det.prime(xs[:W])
        got = [det.slide(xs[W + (t - 1) * S: W + t * S]) for t in range(1, 11)]
        got += det.slides_device(ds.at(W + 10 * S), T - 10)
        st = det.stats()
    assert_same(got, want, f"launch_ray_num={L} {index}")
    assert st["candidates_last"] == min(L, W), st


def test_ablation_work_exceeds_default():
    """No grid filter counts every point: more predicate evaluations than the
    filtered default on the same slides, the same outliers."""
    xs, W, S, Rr, K = STREAMS["walk1d"]()
    T = 20
    case("no grid filter does more work than the default")
    tests = {}
    lists = {}
    for abl in (R.ABL_NONE, R.ABL_NO_FILTER):
        with R.Detector(1, W, S, Rr, K, ablation=abl) as det:
            det.prime(xs[:W])
            lists[abl] = [det.slide(xs[W + (t - 1) * S: W + t * S]) for t in range(1, T + 1)]
            tests[abl] = det.stats()["pair_tests_total"]
    assert_same(lists[R.ABL_NO_FILTER], lists[R.ABL_NONE], "no filter vs default")
    assert tests[R.ABL_NO_FILTER] > 10 * tests[R.ABL_NONE] > 0, tests
