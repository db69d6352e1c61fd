"""GPU tests of batched streams: many independent streams of uneven dim, W,
S, R and K advanced together, each phase one launch over all streams.  Bar: every stream's
outlier list equals the grid oracle's for that stream alone, on every step,
whatever the input path (host or HBM) and whatever the tuning."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import rtod_amd as R
from rtod_amd.synth import calibrate_R

pytestmark = pytest.mark.gpu


def case(name):
    print(f"[case] {name}", flush=True)


def _streams(specs, T, seed0=500):
    """specs: (dim, W, S, K, kind); R calibrated on the first window."""
    descs, data = [], []
    for i, (dim, W, S, K, kind) in enumerate(specs):
        n = W + T * S
        xs = O.gen_walk(seed0 + i, n, dim) if kind == "walk" else O.gen_gmm(seed0 + i, n, dim)
        Rr = calibrate_R(xs[:W], K, seed=i)
        descs.append((dim, W, S, Rr, K))
        data.append(np.ascontiguousarray(xs, np.float32))
    return descs, data


MIXED = [
    (1, 4000, 200, 20, "walk"), (2, 3000, 100, 20, "walk"), (3, 5000, 250, 50, "gmm"),
    (3, 2000, 40, 20, "gmm"), (1, 12000, 3000, 50, "walk"),   # S > one sort tile
    (2, 6000, 300, 50, "gmm"), (1, 999, 1, 5, "walk"),         # S = 1
    (3, 2500, 2500, 20, "gmm"),                                # S = W
    (2, 9000, 4500, 20, "walk"), (3, 7000, 350, 20, "gmm"),
    (1, 2048, 64, 50, "walk"), (3, 3000, 150, 5, "gmm"),
]


def _oracle(descs, data, T):
    return [O.run_stream(xs, W, S, T, Rr, K, "grid") for (dim, W, S, Rr, K), xs in
            zip(descs, data)]


def _check(got_steps, want, name):
    # got_steps[t][i]: stream i at step t+1; want[i][t]
    for t, per in enumerate(got_steps):
        for i, g in enumerate(per):
            w = want[i][t]
            assert np.array_equal(g, w), f"{name}: stream {i} step {t + 1}: " \
                f"gpu-only {sorted(set(g.tolist()) - set(w.tolist()))[:6]} " \
                f"oracle-only {sorted(set(w.tolist()) - set(g.tolist()))[:6]}"


def _run_host(descs, data, T, tuning=None):
    with R.Batch(descs) as b:
        if tuning:
            b.set_tuning(**tuning)
        b.prime([xs[:W] for (dim, W, S, Rr, K), xs in zip(descs, data)])
        out = []
        for t in range(1, T + 1):
            out.append(b.step([xs[W + (t - 1) * S: W + t * S]
                               for (dim, W, S, Rr, K), xs in zip(descs, data)]))
        return out


def _bind(b, descs, data):
    ds = [R.DeviceStream(xs) for xs in data]
    b.bind_device([d.ptr.value for d in ds], [xs.shape[0] for xs in data])
    return ds


def test_batch_mixed_streams_host_input():
    case("batch mixed 12 streams, host input")
    T = 12
    descs, data = _streams(MIXED, T)
    _check(_run_host(descs, data, T), _oracle(descs, data, T), "host input")


@pytest.mark.parametrize("tuning", [
    dict(detect_points_per_wg=16, cell_scale=(1.0, 1.0, 1.0)),
    dict(detect_points_per_wg=256, cell_scale=(0.5, 0.5, 1.0)),
    dict(detect_points_per_wg=4096, cell_scale=(0.25, 0.25, 1.0)),
])
def test_batch_device_input_any_tuning(tuning):
    case(f"batch device input {tuning}")
    T = 10
    descs, data = _streams(MIXED, T)
    want = _oracle(descs, data, T)
    with R.Batch(descs) as b:
        b.set_tuning(**tuning)
        ds = _bind(b, descs, data)
        b.prime_device()
        got = [b.step_device() for _ in range(T)]
        b.close()
        for d in ds:
            d.close()
    _check(got, want, f"device {tuning}")


def test_batch_of_one_equals_detector():
    wl = O.C1
    case("batch of one = detector (C1)")
    xs = np.ascontiguousarray(wl.data(), np.float32)
    T = 30
    desc = [(1, wl.W, wl.S, wl.R, wl.K)]
    got = _run_host(desc, [xs], T)
    want = R.run_stream(xs, wl.W, wl.S, T, wl.R, wl.K)
    for t in range(T):
        assert np.array_equal(got[t][0], want[t]), t


def test_batch_autotune_then_exact():
    case("batch autotune")
    T, tune = 8, 4
    descs, data = _streams(MIXED[:8], T + tune, seed0=900)
    want = _oracle(descs, data, T)
    with R.Batch(descs) as b:
        ds = _bind(b, descs, data)
        chosen = b.autotune(tune)
        assert chosen["us_per_step"] > 0
        assert b.get_tuning()["detect_points_per_wg"] == chosen["detect_points_per_wg"]
        got = [b.step_device() for _ in range(T)]
        st = b.stream_stats(0)
        b.close()
        for d in ds:
            d.close()
    _check(got, want, "after autotune")
    assert st["slides"] == T


def test_batch_errors():
    case("batch state errors")
    descs, data = _streams(MIXED[:2], 2)
    with R.Batch(descs) as b:
        with pytest.raises(R.RtodError):       # step before prime
            b.step([xs[:S] for (dim, W, S, Rr, K), xs in zip(descs, data)])
        with pytest.raises(R.RtodError):       # no device streams bound
            b.step_device()
        with pytest.raises(R.RtodError):       # 3-D cell scale must stay 1
            b.set_tuning(64, (1.0, 1.0, 0.5))


def test_learned_batch_exact_and_trains():
    case("learned batch, 12 streams, device input")
    T = 30
    descs, data = _streams(MIXED, T, seed0=700)
    want = _oracle(descs, data, T)
    with R.Batch(descs, sched_kind=R.SCHED_LEARNED) as b:
        ds = _bind(b, descs, data)
        b.prime_device()
        got = [b.step_device() for _ in range(T)]
        stats = [b.stream_stats(i) for i in range(len(descs))]
        b.close()
        for d in ds:
            d.close()
    _check(got, want, "learned batch")
    # the learned kernels run for the 2-D and 3-D streams (default tuning);
    # 1-D streams, whose near stencil is the whole one, keep the plain ones
    for (dim, *_), st in zip(descs, stats):
        if dim == 1:
            assert st["plain_kernel_slides"] == T and st["predictor_points"] == 0, st
        elif st["candidates_total"] > 0:
            assert st["plain_kernel_slides"] == 0 and st["predictor_points"] > 0, st
    assert max(st["train_steps"] for st in stats) > 0  # the shared predictor trained


def test_learned_batch_learned_dims():
    """learned_dims picks the dimensionalities on the learned kernels: 3-D
    only, then none (every stream on the plain kernels, no training); exact
    in both, and bit 0 (1-D) is refused."""
    case("learned batch, learned_dims 4 and 0")
    T = 12
    descs, data = _streams(MIXED, T, seed0=710)
    want = _oracle(descs, data, T)
    for ld in (4, 0):
        with R.Batch(descs, sched_kind=R.SCHED_LEARNED) as b:
            ds = _bind(b, descs, data)
            b.set_tuning(64, learned_dims=ld)
            assert b.get_tuning()["learned_dims"] == ld
            b.prime_device()
            got = [b.step_device() for _ in range(T)]
            stats = [b.stream_stats(i) for i in range(len(descs))]
            with pytest.raises(R.RtodError):
                b.set_tuning(64, learned_dims=1)
            b.close()
            for d in ds:
                d.close()
        _check(got, want, f"learned batch learned_dims={ld}")
        for (dim, *_), st in zip(descs, stats):
            on = dim > 1 and (ld >> (dim - 1)) & 1
            assert st["plain_kernel_slides"] == (0 if on else T), (dim, ld, st)
        if ld == 0:
            assert max(st["train_steps"] for st in stats) == 0


def test_batch_large_streams_plain_and_learned_exact():
    """256 uneven streams (W up to 100k) over many steps, plain and learned,
    equal the grid oracle on every step.  In a batch, the last detect
    workgroup of a stream publishes the slide (and advances the stream's
    slide counter) while the stream's other detect workgroups may still be
    starting; these take the slide from the counter classify left, else
    they would read the next slide's counters and take a detect ticket,
    making the next slide's epilogue run early and miss outliers."""
    from rtod_amd import synth
    case("batch of 256 streams up to W=100k, plain (40 steps) and learned (20)")
    T = 40
    descs, data = synth.batch_workload(256, seed=11, slides=T)
    cfg = [(d.dim, d.W, d.S, d.R, d.K) for d in descs]
    want = [O.run_stream(xs, d.W, d.S, T, d.R, d.K, "grid", nthreads=16)
            for d, xs in zip(descs, data)]
    for sched, steps in ((R.SCHED_PLAIN, T), (R.SCHED_LEARNED, T // 2)):
        with R.Batch(cfg, sched_kind=sched) as b:
            ds = _bind(b, cfg, data)
            b.prime_device()
            got = [b.step_device() for _ in range(steps)]
            b.close()
            for d in ds:
                d.close()
        _check(got, want, f"256-stream batch sched {sched}")
