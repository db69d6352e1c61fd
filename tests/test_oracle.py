"""Pins the CPU oracle (CPU only).

Every check is chosen so that a plausible mistake fails it: a `<` for `<=`
(lattices with R exactly on a lattice distance), a forgotten self count
(K = count thresholds on lattices), a contracted or double-precision predicate
(exact rounding of the FMA chain), a wrong ring slot (sliding vs from-scratch
windows, golden slot lists), a wrong cell rule (grid vs brute force vs the
independent sorted-window method on clustered, duplicated, clamped and
non-finite data).
"""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


# ---------------------------------------------------------------- predicate
def _round_f32(x: Fraction) -> np.float32:
    """Correct round-to-nearest-even of an exact rational to float32."""
    f = np.float32(float(x))
    best = None
    for c in (np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))):
        err = abs(Fraction(float(c)) - x)
        key = (err, int(np.float32(c).view(np.uint32)) & 1)
        if best is None or key < best[0]:
            best = (key, c)
    return np.float32(best[1])


def _d2_exact(p, q):
    p = [np.float32(v) for v in p]
    q = [np.float32(v) for v in q]
    dx = np.float32(p[0] - q[0])
    d2 = _round_f32(Fraction(float(dx)) ** 2)
    for a in range(1, len(p)):
        dy = np.float32(p[a] - q[a])
        d2 = _round_f32(Fraction(float(dy)) ** 2 + Fraction(float(d2)))
    return d2


def test_predicate_rounds_like_fma_chain():
    rng = np.random.default_rng(1)
    for dim in (1, 2, 3):
        for _ in range(400):
            p = rng.normal(0, 1, dim).astype(np.float32) * np.float32(rng.choice([1e-3, 1, 1e3]))
            q = p + rng.normal(0, 1e-2, dim).astype(np.float32)
            assert O.d2(p, q) == _d2_exact(p, q)


def test_predicate_fma_differs_from_double_sometimes():
    """The contract is the FP32 FMA chain, not a double sum rounded once:
    find inputs where the two differ and check the oracle takes the chain."""
    rng = np.random.default_rng(2)
    seen = 0
    for _ in range(20000):
        p = rng.uniform(-1, 1, 3).astype(np.float32)
        q = rng.uniform(-1, 1, 3).astype(np.float32)
        dbl = np.float32(sum((float(a) - float(b)) ** 2 for a, b in zip(p, q)))
        ex = _d2_exact(p, q)
        if dbl != ex:
            seen += 1
            assert O.d2(p, q) == ex
    assert seen > 0


# ------------------------------------------------------------ closed forms
def test_lattice_1d_counts_and_le():
    h, W = 0.25, 64
    win = (np.arange(W) * h).astype(np.float32).reshape(-1, 1)
    out, cnt = O.brute(win, 0.5, 100, counts=True)
    # <= R: j in i-2..i+2 (distance 0.5 == R counts); self included
    want = np.minimum(np.arange(W), 2) + np.minimum(W - 1 - np.arange(W), 2) + 1
    assert np.array_equal(cnt, want)
    assert O.brute(win, 0.5, 4).tolist() == [0, 1, W - 2, W - 1]
    assert len(O.brute(win, 0.5, 5)) == W
    # the reference's strict FP64 rule |dx| < R: only i-1..i+1
    _, cnt_ref = O.brute(win, 0.5, 100, semantics=1, counts=True)
    want_ref = np.minimum(np.arange(W), 1) + np.minimum(W - 1 - np.arange(W), 1) + 1
    assert np.array_equal(cnt_ref, want_ref)


def test_lattice_3d_cube():
    n = 6
    g = np.stack(np.meshgrid(*(np.arange(n),) * 3, indexing="ij"), -1).reshape(-1, 3)
    win = g.astype(np.float32)
    # R = 1: self + face neighbours (diagonals are sqrt(2) > 1)
    _, cnt = O.brute(win, 1.0, 100, counts=True)
    want = 1 + sum(((g[:, a] > 0).astype(int) + (g[:, a] < n - 1).astype(int)) for a in range(3))
    assert np.array_equal(cnt, want)
    interior = np.all((g > 0) & (g < n - 1), axis=1)
    assert np.array_equal(O.brute(win, 1.0, 6), np.nonzero(~interior)[0])


def test_isolated_dense_and_k_ge_w():
    rng = np.random.default_rng(3)
    cluster = np.zeros((60, 2), np.float32)
    far = rng.uniform(100, 200, (5, 2)).astype(np.float32)
    win = np.concatenate([cluster, far])
    out = O.brute(win, 0.1, 50)
    assert out.tolist() == list(range(60, 65))  # cluster of 60 >= K+1 identical points
    assert len(O.brute(win, 0.1, 65)) == 65     # K >= W: everything


# ------------------------------------------------- independent method agree
@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("R,K", [(0.1, 5), (0.45, 50), (1.0, 200)])
def test_sorted1d_equals_brute(seed, R, K):
    win = O.gen_walk(seed, 3000, 1)
    a, ca = O.brute(win, R, K, counts=True)
    b, cb = O.sorted1d(win, R, K, counts=True)
    assert np.array_equal(a, b)
    assert np.array_equal(np.minimum(ca, K + 1), cb)


def _grid_vs_brute(xs, W, S, T, R, K, nthreads=1):
    g = O.run_stream(xs, W, S, T, R, K, "grid", nthreads=nthreads)
    b = O.run_stream(xs, W, S, T, R, K, "brute")
    for t, (x, y) in enumerate(zip(g, b), 1):
        assert np.array_equal(x, y), f"slide {t}"


@pytest.mark.parametrize("dim,kind,R,K", [(1, "walk", 0.45, 10), (2, "walk", 0.3, 20),
                                          (3, "walk", 0.3, 20), (2, "gmm", 1.5, 20),
                                          (3, "gmm", 1.9, 50)])
def test_grid_equals_brute(dim, kind, R, K):
    W, S, T = 2400, 120, 25
    xs = O.gen_walk(7, W + T * S, dim) if kind == "walk" else O.gen_gmm(8, W + T * S, dim)
    _grid_vs_brute(xs, W, S, T, R, K)
    _grid_vs_brute(xs, W, S, T, R, K, nthreads=4)


def test_grid_edge_data():
    rng = np.random.default_rng(4)
    W, S, T = 900, 90, 8
    xs = rng.integers(0, 12, (W + T * S, 2)).astype(np.float32) * np.float32(0.1)
    xs[::50] += np.float32(5e9)           # clamped cells
    xs[7] = [np.nan, 0.0]                 # non-finite
    xs[333] = [np.inf, np.inf]
    _grid_vs_brute(xs, W, S, T, 0.15, 12)
    _grid_vs_brute(xs, W, S, T, 0.15, 1000)


def test_grid_shapes():
    xs = O.gen_walk(9, 6000, 1)
    for W, S in ((1000, 1000), (999, 1), (1000, 500)):
        T = min(20, (xs.shape[0] - W) // S)
        _grid_vs_brute(xs, W, S, T, 0.45, 8)


def test_grid_invalid_configs():
    for args in ((0, 100, 10, 0.5, 5), (1, 100, 30, 0.5, 5), (1, 100, 10, 0.0, 5),
                 (1, 100, 10, 0.5, 0), (4, 100, 10, 0.5, 5)):
        with pytest.raises(ValueError):
            O.Grid(*args)


# -------------------------------------------------------------- invariants
def test_permutation_invariance():
    rng = np.random.default_rng(5)
    win = O.gen_gmm(10, 2000, 3)
    perm = rng.permutation(2000)
    a = O.brute(win, 1.9, 30)
    b = O.brute(win[perm], 1.9, 30)
    assert np.array_equal(np.sort(perm[b]), a)


def test_monotone_in_r_and_k():
    win = O.gen_walk(11, 4000, 1)
    prev = None
    for R in (0.1, 0.2, 0.4, 0.8):
        s = set(O.brute(win, R, 40).tolist())
        assert prev is None or s <= prev
        prev = s
    prev = None
    for K in (5, 20, 40, 80):
        s = set(O.brute(win, 0.3, K).tolist())
        assert prev is None or prev <= s
        prev = s


def test_cells_have_bounded_diameter():
    """Two points in one unclamped cell always pass the predicate, also at
    large magnitudes and next to cell boundaries: the dense rule is exact."""
    rng = np.random.default_rng(12)
    for dim, R in ((1, 0.45), (2, 0.3), (3, 1.9)):
        c = R * (1 - 1e-5) / np.sqrt(dim)
        for scale in (1.0, 1e3, 1e5):
            k = rng.integers(-1000, 1000, (400, dim)) * scale
            base = (k * c).astype(np.float64)
            p = (base + rng.uniform(-1e-6, 1, (400, dim)) * c).astype(np.float32)
            q = (base + rng.uniform(0, 1 + 1e-6, (400, dim)) * c).astype(np.float32)
            checked = 0
            for a, b in zip(p, q):
                key = O.cell_key(a, R)
                if key == O.cell_key(b, R) and _unclamped(key, dim):
                    assert O.d2(a, b) <= np.float32(R) * np.float32(R)
                    checked += 1
            assert checked > 0 or scale * 1000 * c > 2 ** (20 if dim == 3 else 30) * c


def _unclamped(key, dim):
    F = 21 if dim == 3 else 31
    M = (1 << F) - 1
    return all(0 < (key >> (F * (dim - 1 - a))) & M < M for a in range(dim))


def test_cell_key_packing():
    R = 1.0
    c = (1 - 1e-5)
    assert O.cell_key(np.float32([0.5]), R) == 2 ** 30
    assert O.cell_key(np.float32([-0.5]), R) == 2 ** 30 - 1
    assert O.cell_key(np.float32([2.5 * c]), R) == 2 ** 30 + 2
    k3 = O.cell_key(np.float32([0.1, 0.1, -0.1]), R * np.sqrt(3))
    assert k3 == (2 ** 20 << 42) | (2 ** 20 << 21) | (2 ** 20 - 1)
    big = O.cell_key(np.float32([3e38]), 1.0)
    assert big == 2 ** 31 - 1  # clamped to the top field value


# ----------------------------------------------------------------- goldens
def _golden():
    with open(os.path.join(HERE, "golden", "golden.json")) as f:
        return json.load(f)


def test_generators_match_golden():
    import hashlib

    g = _golden()["generators"]
    for wl in (O.C1, O.C2, O.C3):
        xs = wl.data()
        assert hashlib.sha256(xs.tobytes()).hexdigest() == g[wl.name]["sha256"], wl.name
        assert xs.shape == (wl.N, wl.dim)


def test_streams_match_golden():
    gs = _golden()["streams"]
    for wl in (O.C1, O.C3):
        gd = gs[wl.name]
        xs = wl.data()
        got = O.run_stream(xs, wl.W, wl.S, gd["slides"], wl.R, wl.K, "grid")
        for t, (a, b) in enumerate(zip(got, gd["outliers"]), 1):
            assert a.tolist() == b, f"{wl.name} slide {t}"


def test_workload_constants():
    assert O.C1.T == 380 and O.C2.T == 189 and O.C3.T == 1130
    assert (O.C2.W, O.C2.S, O.C2.K, O.C2.R) == (100_000, 5_000, 50, 0.45)


# ---------------------------------------------------------------------------
# predictor MLP reference (learned scheduler)
# ---------------------------------------------------------------------------
def _bf16_rne_exact(x: float) -> float:
    """Round-to-nearest-even onto the bfloat16 grid, by exact rationals."""
    if x == 0.0 or not np.isfinite(x):
        return x
    s = -1.0 if x < 0 else 1.0
    a = Fraction(abs(x))
    e = 0
    while a >= 2:
        a /= 2
        e += 1
    while a < 1:
        a *= 2
        e -= 1
    scaled = a * 128            # 7 explicit mantissa bits: grid step 1/128
    lo = scaled.numerator // scaled.denominator
    rem = scaled - lo
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and lo % 2 == 1):
        lo += 1
    return s * float(Fraction(lo, 128) * Fraction(2) ** e)


def test_bf16_round_is_nearest_even():
    f32 = lambda v: float(np.float32(v))
    ties = [1 + 2 ** -8, 1 + 3 * 2 ** -8, -(1 + 2 ** -8), 3 * 2 ** -9 * (1 + 2 ** -8)]
    assert O.bf16_round(ties[0]) == 1.0                    # tie -> even (down)
    assert O.bf16_round(ties[1]) == 1 + 2 ** -6            # tie -> even (up)
    assert O.bf16_round(f32(1 + 2 ** -8 + 2 ** -20)) == 1 + 2 ** -7
    rng = np.random.default_rng(3)
    vals = list(rng.normal(0, 10, 2000).astype(np.float32)) + ties
    vals += list((rng.integers(-(2 ** 20), 2 ** 20, 500) * 2.0 ** -16).astype(np.float32))
    for v in vals:
        assert O.bf16_round(float(v)) == _bf16_rne_exact(f32(v)), v
    assert np.isnan(O.bf16_round(float("nan")))
    assert O.bf16_round(float("inf")) == float("inf")


def _rand_params(rng, scale=0.4):
    return rng.normal(0, scale, O.MLP_NP).astype(np.float32)


def test_mlp_gradient_matches_central_differences():
    rng = np.random.default_rng(11)
    P = _rand_params(rng)
    X = rng.uniform(-0.5, 1.0, (48, 6)).astype(np.float32)
    y = rng.uniform(0, 1, 48).astype(np.float32)
    loss, g = O.mlp_grad(P, X, y, bf16=False)
    assert abs(loss - O.mlp_loss64(P.astype(np.float64), X, y)) < 1e-6
    h = 1e-5
    P64 = P.astype(np.float64)
    fd = np.empty(O.MLP_NP)
    for i in range(O.MLP_NP):
        Pp = P64.copy(); Pp[i] += h
        Pm = P64.copy(); Pm[i] -= h
        fd[i] = (O.mlp_loss64(Pp, X, y) - O.mlp_loss64(Pm, X, y)) / (2 * h)
    err = np.abs(g - fd)
    tol = 1e-4 * np.abs(fd).max() + 1e-3 * np.abs(fd)
    # a ReLU kink within h of a sample may spoil a handful of differences
    assert np.count_nonzero(err > tol) <= 3
    assert np.abs(fd).max() > 1e-3


def test_mlp_closed_form_constant_network():
    # all weights zero, output bias z0: yh = sigmoid(z0) for every input;
    # the only non-zero gradient entry is b3 = mean(2 e yh (1-yh))
    z0 = 0.3
    P = np.zeros(O.MLP_NP, np.float32)
    P[-1] = z0
    rng = np.random.default_rng(5)
    X = rng.uniform(-1, 1, (20, 6)).astype(np.float32)
    y = rng.uniform(0, 1, 20).astype(np.float32)
    yh = 1 / (1 + np.exp(-z0))
    assert np.allclose(O.mlp_forward(P, X), yh, rtol=1e-6)
    loss, g = O.mlp_grad(P, X, y)
    assert np.isclose(loss, np.mean((yh - y) ** 2), rtol=1e-5)
    assert np.isclose(g[-1], np.mean(2 * (yh - y) * yh * (1 - yh)), rtol=1e-5)
    assert np.count_nonzero(g[:-1]) == 0


def test_mlp_bf16_emulation_close_to_fp32():
    rng = np.random.default_rng(8)
    P = _rand_params(rng)
    X = rng.uniform(-0.5, 1.0, (128, 6)).astype(np.float32)
    y = rng.uniform(0, 1, 128).astype(np.float32)
    a = O.mlp_forward(P, X, bf16=True)
    b = O.mlp_forward(P, X, bf16=False)
    assert np.abs(a - b).max() < 2e-2 and np.abs(a - b).max() > 0
    _, ga = O.mlp_grad(P, X, y, bf16=True)
    _, gb = O.mlp_grad(P, X, y, bf16=False)
    assert 0 < np.abs(ga - gb).max() < 0.25 * np.abs(gb).max()
    assert np.dot(ga, gb) / (np.linalg.norm(ga) * np.linalg.norm(gb)) > 0.99


# ------------------------------------------- contract vs reference semantics
def _status_diff_brute(win, R, K):
    """Points whose outlier status differs between the contract and the
    reference's FP64 strict rule, by brute force under both."""
    a = set(O.brute(win, R, K, semantics=0).tolist())
    b = set(O.brute(win, R, K, semantics=1).tolist())
    return len(a ^ b)


def test_semantics_diff_lattice_closed_form():
    """x_i = i/4, R = 1/2 exactly: the contract counts the two lattice points
    at distance exactly R, the reference's strict test does not.  Interior
    points: 5 vs 3 neighbours (K = 4 splits them), the two points at each end
    are outliers either way; every count differs; 2 (W - 2) ordered pairs sit
    at distance exactly R."""
    W = 256
    win = (np.arange(W, dtype=np.float64) * 0.25).astype(np.float32).reshape(-1, 1)
    got = O.semantics_diff(win, 0.5, 4)
    assert got == {"status": W - 4, "counts": W, "pairs": 2 * (W - 2)}
    assert got["status"] == _status_diff_brute(win, 0.5, 4)
    # 3-D integer lattice, R = 1: the six axis neighbours sit exactly at R
    g = np.arange(6, dtype=np.float32)
    cube = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    got3 = O.semantics_diff(cube, 1.0, 3)
    assert got3["status"] == _status_diff_brute(cube, 1.0, 3)
    inner = sum(1 for p in cube for a in range(3) for s in (-1, 1)
                if 0 <= p[a] + s <= 5)  # ordered pairs at distance exactly 1
    assert got3["pairs"] == inner


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_semantics_diff_equals_brute(dim):
    """Windows with many exact ties (values on a 1/8 grid, R a grid distance)
    and continuous ones: the fast counter's status differences equal the
    symmetric difference of the two brute-force outlier sets."""
    rng = np.random.default_rng(100 + dim)
    for trial in range(6):
        W = int(rng.integers(300, 1500))
        if trial % 2 == 0:
            win = (rng.integers(0, 40, (W, dim)) / 8.0).astype(np.float32)
            R = float(rng.choice([0.5, 0.625, 1.0]))
        else:
            win = rng.normal(0, 3, (W, dim)).astype(np.float32)
            R = float(rng.uniform(0.2, 1.5))
        K = int(rng.integers(1, 40))
        assert O.semantics_diff(win, R, K)["status"] == _status_diff_brute(win, R, K), (dim, trial)


def test_semantics_diff_on_configs():
    """The measured 1-D and 3-D streams, every slide: how often the build's
    FP32 `d2 <= R*R` and the reference's FP64 `|dx| < R` / `sqdist < R2`
    (optixScan.cu:170-189) give a different outlier set.  C1 and C2: never
    (C2 has pairs on which the rules differ, none of which moves a count
    across K).  C3: two points, in two slides; those slides are re-checked by
    brute force under both rules."""
    want = {"walk1d-W10k-S500": 0, "walk1d-W100k-S5k": 0, "gmm3d-W10k-S500": 2}
    for wl in (O.C1, O.C2, O.C3):
        xs = wl.data()
        tot, per = O.semantics_diff_stream(xs, wl.W, wl.S, wl.T, wl.R, wl.K)
        assert tot["slides_checked"] == wl.T
        assert tot["status"] == want[wl.name], (wl.name, tot)
        for t in np.nonzero(per > 0)[0][:2] + 1:
            idx = np.arange(t * wl.S, t * wl.S + wl.W)
            win = np.empty((wl.W, wl.dim), np.float32)
            win[idx % wl.W] = xs[idx]
            assert _status_diff_brute(win, wl.R, wl.K) == per[t - 1], (wl.name, t)


# ------------------------------------------------------------------ hits
def test_grid_hits_closed_form_and_brute():
    """Hits (the sum over candidates of min(count, K + 1)).  K >= W makes
    every point a candidate whose count stops nowhere, so hits is the sum of
    the full counts: on the lattice x_i = i/4 with R = 1/2 that is 5W - 6
    (interior points 5, the two points at each end 3 and 4); on random
    windows the sum of brute-force counts.  With small K each candidate adds
    min(count, K + 1): bounded by (K + 1) * candidates."""
    W, S = 256, 64
    xs = (np.arange(W + S, dtype=np.float64) * 0.25).astype(np.float32).reshape(-1, 1)
    g = O.Grid(1, W, S, 0.5, W)
    g.prime(xs[:W])
    g.slide(xs[W:W + S])
    assert g.stats()["hits"] == 5 * W - 6
    rng = np.random.default_rng(5)
    for dim in (1, 2, 3):
        W, S = 600, 100
        xs = rng.normal(0, 2, (W + S, dim)).astype(np.float32)
        g = O.Grid(dim, W, S, 0.9, W)
        g.prime(xs[:W])
        g.slide(xs[W:W + S])
        win = np.empty((W, dim), np.float32)
        idx = np.arange(S, S + W)
        win[idx % W] = xs[idx]
        _, cnt = O.brute(win, 0.9, W, counts=True)
        assert g.stats()["hits"] == int(cnt.sum()), dim
        g2 = O.Grid(dim, W, S, 0.9, 5)
        g2.prime(xs[:W])
        g2.slide(xs[W:W + S])
        st = g2.stats()
        assert 0 < st["hits"] <= 6 * st["candidates"]
