"""GPU tests of the learned work scheduler.

* The predictor's forward and backward kernels (BF16 MFMA, FP32 accumulation)
  against the CPU MLP reference with the same BF16 operand rounding
  (oracle/rtod_oracle.c), which is itself pinned by central differences.
* Exactness: with the learned schedule the outlier set equals the grid
  oracle's on every slide (the schedule only reorders exact counts).
* The fallback rule and the predictor API.
* Gain: after online training, the learned schedule evaluates fewer pair
  predicates than the plain schedule and than the same method without a
  trained network (untrained, frozen).
"""
import math

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


def _features(rng, n):
    X = np.empty((n, 6), np.float32)
    X[:, :3] = rng.integers(0, 52, (n, 3)) / 51.0
    X[:, 3:] = rng.uniform(-0.5, 0.5, (n, 3))
    return X


def test_mlp_forward_matches_oracle():
    case("mlp forward n=1000")
    rng = np.random.default_rng(1)
    P = rng.normal(0, 0.4, R.MLP_NPARAMS).astype(np.float32)
    X = _features(rng, 1000)
    got = R.mlp_forward(P, X)
    want16 = O.mlp_forward(P, X, bf16=True)
    want32 = O.mlp_forward(P, X, bf16=False)
    err16 = np.abs(got - want16)
    # same operand rounding; only the FP32 accumulation order differs
    assert err16.max() < 2e-3, err16.max()
    assert err16.mean() < 0.25 * np.abs(got - want32).mean()


def test_mlp_bf16_rounding_is_nearest_even_on_device():
    case("mlp bf16 rounding probe")
    # x0 -> h1[0] = relu(bf16(x0)) -> h2[0] = h1[0] -> z = h2[0]: y = sigmoid(bf16(x0))
    P = np.zeros(R.MLP_NPARAMS, np.float32)
    P[0 * 32 + 0] = 1.0            # W1[0][0]
    P[224 + 0 * 32 + 0] = 1.0      # W2[0][0]
    P[1280 + 0] = 1.0              # W3[0]
    xs = np.array([1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -8 + 2 ** -15, 2.0 + 2 ** -7,
                   0.75 + 2 ** -10], np.float32)
    X = np.zeros((len(xs), 6), np.float32)
    X[:, 0] = xs
    got = R.mlp_forward(P, X)
    for x, g in zip(xs, got):
        q = O.bf16_round(float(x))
        assert abs(g - 1 / (1 + math.exp(-q))) < 2e-7, (x, q, g)


@pytest.mark.parametrize("n", [128, 77])
def test_mlp_grad_matches_oracle(n):
    case(f"mlp grad n={n}")
    rng = np.random.default_rng(n)
    P = rng.normal(0, 0.4, R.MLP_NPARAMS).astype(np.float32)
    X = _features(rng, n)
    y = np.where(rng.uniform(size=n) < 0.6, 1.0, rng.uniform(0, 0.5, n)).astype(np.float32)
    loss, g = R.mlp_grad(P, X, y)
    loss16, g16 = O.mlp_grad(P, X, y, bf16=True)
    loss32, g32 = O.mlp_grad(P, X, y, bf16=False)
    scale = np.abs(g16).max()
    assert abs(loss - loss16) < 1e-3 * max(loss16, 1e-3)
    assert np.abs(g - g16).max() < 2e-2 * scale, np.abs(g - g16).max() / scale
    # against plain FP32 the BF16 operand rounding moves single entries by up
    # to ~20% of the largest (the CPU reference shows the same gap); the
    # direction stays
    cos = float(np.dot(g, g32) / (np.linalg.norm(g) * np.linalg.norm(g32)))
    assert cos > 0.99, cos
    # a small step along -g lowers the exact (FP64) loss
    P2 = (P - 0.05 * g / np.linalg.norm(g)).astype(np.float64)
    assert O.mlp_loss64(P2, X, y) < O.mlp_loss64(P.astype(np.float64), X, y)


ALW = R.FLAG_LEARN_ALWAYS  # the learned kernels on every slide (cost rule off)


def _learned_stream(xs, W, S, T, Rr, K, params=None, frozen=False, **kw):
    """Slides 1..T with the learned schedule, its kernels on every slide
    unless flags say otherwise; returns (lists, stats)."""
    xs = np.ascontiguousarray(xs, np.float32)
    kw.setdefault("flags", ALW)
    with R.Detector(xs.shape[1], W, S, Rr, K, sched_kind=R.SCHED_LEARNED, **kw) as det:
        if params is not None:
            det.set_predictor(params, frozen=frozen)
        det.prime(xs[:W])
        out = [det.slide(xs[W + (t - 1) * S: W + t * S]) for t in range(1, T + 1)]
        return out, det.stats()


def test_learned_exact_c3_vs_grid():
    wl = O.C3
    case(wl.name + " learned")
    xs = wl.data()
    T = 150
    got, st = _learned_stream(xs, wl.W, wl.S, T, wl.R, wl.K)
    assert_same(got, O.run_stream(xs, wl.W, wl.S, T, wl.R, wl.K, "grid"), "C3 learned vs grid")
    # >= 1 step per slide once replay data exists (reported up to a slide late)
    assert st["train_steps"] >= T - 3
    assert 0 < st["predictor_points"] <= st["candidates_total"]   # cell units
    print(f"C3 learned: misses {st['predictor_misses'] / st['predictor_points']:.3f} "
          f"fallback slides {st['fallbacks']} steps {st['train_steps']} "
          f"loss {st['train_loss_last']:.4f}", flush=True)


@pytest.mark.parametrize("name,dim,gen,W,S,T,Rr,K", [
    ("walk1d", 1, "walk", 10_000, 500, 60, 0.45, 50),
    ("walk2d", 2, "walk", 6000, 300, 40, 0.3, 20),
    ("gmm2d", 2, "gmm", 6000, 300, 40, 1.5, 30),
    ("gmm3d-smallK", 3, "gmm", 4000, 200, 40, 2.5, 3),
    ("gmm3d-bigK", 3, "gmm", 4000, 400, 30, 3.0, 300),   # features from HBM, not LDS
])
def test_learned_exact_vs_grid(name, dim, gen, W, S, T, Rr, K):
    case(f"learned {name}")
    n = W + T * S
    xs = O.gen_walk(31, n, dim) if gen == "walk" else O.gen_gmm(32, n, dim)
    got, _ = _learned_stream(xs, W, S, T, Rr, K)
    assert_same(got, O.run_stream(xs, W, S, T, Rr, K, "grid"), name)


def test_learned_fallback_rule():
    case("learned fallback thresholds")
    W, S, T = 4000, 200, 25
    xs = O.gen_gmm(41, W + T * S, 3)
    want = O.run_stream(xs, W, S, T, 2.0, 20, "grid")
    # threshold 1: every slide after the first two (slide t follows the
    # decision slide t - 2 made) is scheduled by prediction
    got, st = _learned_stream(xs, W, S, T, 2.0, 20, miss_threshold=1.0)
    assert_same(got, want, "threshold 1")
    assert st["fallbacks"] == 2 and st["sched_kind"] == R.SCHED_LEARNED
    # a tiny threshold with a frozen untrained network: never engages
    p0 = np.zeros(R.MLP_NPARAMS, np.float32)
    p0[-1] = 4.0  # every candidate a probable inlier: many misses
    got, st = _learned_stream(xs, W, S, T, 2.0, 20, params=p0, frozen=True,
                              miss_threshold=1e-6)
    assert_same(got, want, "threshold ~0")
    assert st["fallbacks"] == T and st["sched_kind"] == R.SCHED_PLAIN
    assert st["train_steps"] == 0
    assert st["predictor_misses"] > 0


def test_predictor_api_roundtrip_and_freeze():
    case("predictor get/set/freeze")
    W, S, T = 3000, 150, 10
    xs = O.gen_gmm(43, W + T * S, 3)
    rng = np.random.default_rng(2)
    p = rng.normal(0, 0.3, R.MLP_NPARAMS).astype(np.float32)
    with R.Detector(3, W, S, 2.0, 20, sched_kind=R.SCHED_LEARNED, flags=ALW) as det:
        det.set_predictor(p, frozen=True)
        assert np.array_equal(det.get_predictor(), p)
        det.prime(xs[:W])
        for t in range(1, T + 1):
            det.slide(xs[W + (t - 1) * S: W + t * S])
        assert np.array_equal(det.get_predictor(), p)      # frozen: unchanged
        det.set_predictor(p, frozen=False)
        det.prime(xs[:W])
        for t in range(1, T + 1):
            det.slide(xs[W + (t - 1) * S: W + t * S])
        assert not np.array_equal(det.get_predictor(), p)  # trained
        with pytest.raises(R.RtodError):
            det.set_predictor(p[:-1])
    with R.Detector(3, W, S, 2.0, 20) as plain, pytest.raises(R.RtodError):
        plain.get_predictor()


def test_learned_phase_api_equals_slide():
    case("learned phase API")
    W, S, T = 3000, 150, 12
    xs = O.gen_gmm(44, W + T * S, 3)
    want = O.run_stream(xs, W, S, T, 2.0, 20, "grid")
    with R.Detector(3, W, S, 2.0, 20, sched_kind=R.SCHED_LEARNED, flags=ALW) as det:
        det.prime(xs[:W])
        got = []
        for t in range(1, T + 1):
            det.copy_new_points(xs[W + (t - 1) * S: W + t * S])
            det.prepare_cell()
            det.build_index()
            det.detect_outlier()
            got.append(det.copy_outlier())
    assert_same(got, want, "learned phase API")


def test_trained_predictor_gain():
    """What the trained network buys: fewer predicate evaluations than the
    plain schedule and than the untrained network, same outliers."""
    wl = O.C3
    case(wl.name + " learned gain")
    xs = np.ascontiguousarray(wl.data(), np.float32)
    W, S = wl.W, wl.S
    T_train, T_eval = 200, 60
    want = O.run_stream(xs, W, S, T_train + T_eval, wl.R, wl.K, "grid")[T_train:]

    def evaluate(det, start):
        lists, tests, t0 = [], 0, det.stats()["pair_tests_total"]
        for t in range(start, start + T_eval):
            lists.append(det.slide(xs[W + (t - 1) * S: W + t * S]))
        return lists, det.stats()["pair_tests_total"] - t0, det.stats()

    # trained online over T_train slides, then measured (still training)
    with R.Detector(3, W, S, wl.R, wl.K, sched_kind=R.SCHED_LEARNED,
                    flags=ALW) as det:
        det.prime(xs[:W])
        for t in range(1, T_train + 1):
            det.slide(xs[W + (t - 1) * S: W + t * S])
        m0, p0, f0 = (det.stats()[k] for k in ("predictor_misses", "predictor_points",
                                                "fallbacks"))
        got, tests_learned, st = evaluate(det, T_train + 1)
        miss_rate = (st["predictor_misses"] - m0) / (st["predictor_points"] - p0)
        fallback_eval = st["fallbacks"] - f0
    assert_same(got, want, "trained")
    # the same slides, plain schedule
    with R.Detector(3, W, S, wl.R, wl.K) as det:
        det.prime(xs[:W])
        for t in range(1, T_train + 1):
            det.slide(xs[W + (t - 1) * S: W + t * S])
        got, tests_plain, _ = evaluate(det, T_train + 1)
    assert_same(got, want, "plain")
    # the untrained network, frozen (the method without training)
    with R.Detector(3, W, S, wl.R, wl.K, sched_kind=R.SCHED_LEARNED,
                    flags=ALW) as det:
        det.set_predictor(det.get_predictor(), frozen=True)
        det.prime(xs[:W])
        for t in range(1, T_train + 1):
            det.slide(xs[W + (t - 1) * S: W + t * S])
        got, tests_untrained, _ = evaluate(det, T_train + 1)
    assert_same(got, want, "untrained")
    print(f"C3 pair tests over {T_eval} slides: plain {tests_plain} untrained "
          f"{tests_untrained} trained {tests_learned} (miss rate {miss_rate:.3f}, "
          f"fallback slides {fallback_eval})", flush=True)
    assert miss_rate < 0.25
    assert fallback_eval <= T_eval // 10
    assert tests_learned < 0.9 * tests_plain
    assert tests_learned < 0.9 * tests_untrained


def _plain_stats(xs, W, S, T, Rr, K):
    with R.Detector(xs.shape[1], W, S, Rr, K) as det:
        det.prime(xs[:W])
        for t in range(1, T + 1):
            det.slide(xs[W + (t - 1) * S: W + t * S])
        return det.stats()


@pytest.mark.parametrize("wl,T", [(O.C1, 40), (O.C3, 30)], ids=["C1", "C3"])
def test_cost_rule_runs_plain_kernels(wl, T):
    """The learned schedule's cost rule: in 1-D (a unit's near stencil is its
    whole stencil) and with no more cell units than the detect's resident
    wavefronts (C3: about 2.7k units, each with a wavefront of its own), the
    prediction cannot shorten a slide, so every slide runs the plain kernels:
    no inference and no training, the plain schedule's outliers and pair
    tests, every slide counted as a fallback."""
    case(f"{wl.name} learned, cost rule")
    xs = wl.data(wl.W + T * wl.S)
    got, st = _learned_stream(xs, wl.W, wl.S, T, wl.R, wl.K, flags=0)
    assert_same(got, O.run_stream(xs, wl.W, wl.S, T, wl.R, wl.K, "grid"), wl.name)
    assert st["plain_kernel_slides"] == T and st["fallbacks"] == T, st
    assert st["predictor_points"] == 0 and st["train_steps"] == 0, st
    assert st["sched_kind"] == R.SCHED_PLAIN
    assert st["pair_tests_total"] == _plain_stats(xs, wl.W, wl.S, T, wl.R, wl.K)["pair_tests_total"]


def test_cost_rule_many_units_runs_learned_kernels():
    """3-D W=100k, S=5k (about 25k cell units, more than the detect's resident
    wavefronts): the first slide, before any unit count is known, runs the
    plain kernels and every later one the learned kernels -- one per call,
    then pipelined (the stage launches switch from one per slide to two
    within the run), all exact against the grid oracle."""
    from rtod_amd import synth
    wl = synth.WORKLOADS["gmm3d-W100k-S5k"]
    T1, T2 = 6, 12
    T = T1 + T2
    case(f"{wl.name} learned, cost rule, {T1} single + {T2} pipelined slides")
    xs = np.ascontiguousarray(wl.data(wl.W + T * wl.S), np.float32)
    want = O.run_stream(xs, wl.W, wl.S, T, wl.R, wl.K, "grid", nthreads=8)
    with R.Detector(3, wl.W, wl.S, wl.R, wl.K, sched_kind=R.SCHED_LEARNED) as det:
        det.prime(xs[:wl.W])
        got = [det.slide(xs[wl.W + (t - 1) * wl.S: wl.W + t * wl.S]) for t in range(1, T1 + 1)]
        st1 = det.stats()
        det.prime(xs[:wl.W])  # the unit count of the last slide is kept
        got2 = det.slides(xs[wl.W: wl.W + T2 * wl.S], T2)
        st2 = det.stats()
    assert_same(got, want[:T1], "one per call")
    assert_same(got2, want[:T2], "pipelined")
    assert st1["plain_kernel_slides"] == 1, st1
    # pipelined: the decision of slide i reads the last collected slide's
    # units, known from before the prime
    assert st2["plain_kernel_slides"] == 1, st2
    assert st2["predictor_points"] > 6144 * (T1 - 1 + T2), st2
    assert st2["train_steps"] > 0


def test_learned_time_vs_plain_large_3d():
    """Time as well as work, where the learned kernels run (the 3-D W=100k
    stream, ~34k cell units a slide): after 200 online slides the learned
    schedule does under 60% of the plain schedule's predicate evaluations
    and is not slower (per-call device time, the mean over 60 slides; the
    bench measures it about 5% faster)."""
    from rtod_amd import synth
    wl = synth.WORKLOADS["gmm3d-W100k-S5k"]
    T_train, T_eval = 200, 60
    case(f"{wl.name} learned vs plain time, {T_train} + {T_eval} slides")
    xs = np.ascontiguousarray(wl.data(wl.W + (T_train + T_eval) * wl.S), np.float32)

    def run(**kw):
        with R.Detector(3, wl.W, wl.S, wl.R, wl.K, **kw) as det:
            det.prime(xs[:wl.W])
            lists, us = [], []
            for t in range(1, T_train + T_eval + 1):
                o = det.slide(xs[wl.W + (t - 1) * wl.S: wl.W + t * wl.S])
                if t > T_train:
                    lists.append(o)
                    us.append(det.times.total * 1e3)
            st = det.stats()
        return lists, float(np.mean(us)), st

    got_p, us_p, st_p = run()
    got_l, us_l, st_l = run(sched_kind=R.SCHED_LEARNED)
    assert_same(got_l, got_p, "learned vs plain")
    assert st_l["plain_kernel_slides"] == 1, st_l  # the first slide only
    print(f"{wl.name}: plain {us_p:.1f} us, learned {us_l:.1f} us per slide "
          f"({us_p / us_l:.3f}x); pair tests {st_p['pair_tests_total']} -> "
          f"{st_l['pair_tests_total']}", flush=True)
    assert st_l["pair_tests_total"] < 0.6 * st_p["pair_tests_total"]
    assert us_l <= 1.03 * us_p, (us_l, us_p)
