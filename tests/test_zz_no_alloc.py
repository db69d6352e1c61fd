"""rtod_mlp_grad_static (the gradient entry point that allocates nothing:
its inputs and results are staged in the library's fixed device scratch
area) against rtod_mlp_grad (per-call allocations, the same kernel), two
calls in a row, the second with fewer samples, then the size check."""
import numpy as np
import pytest

import rtod_amd as R

pytestmark = pytest.mark.gpu


def _inputs(seed, n):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, R.MLP_NPARAMS).astype(np.float32)
    X = rng.uniform(-1, 1, (n, 6)).astype(np.float32)
    y = rng.uniform(0, 1, n).astype(np.float32)
    return p, X, y


def _close(a, b, what):
    # the same kernel on the same inputs: only the order of its LDS
    # accumulations (atomics) can differ, a few ulps of the largest entry
    tol = 1e-5 * max(1e-30, float(np.max(np.abs(b))))
    assert np.max(np.abs(a - b)) <= tol, (what, float(np.max(np.abs(a - b))), tol)


def test_static_grad_equals_allocating_grad():
    print("[case] rtod_mlp_grad_static vs rtod_mlp_grad, n=128 then n=37", flush=True)
    for seed, n in ((1, 128), (2, 37)):
        p, X, y = _inputs(seed, n)
        l0, g0 = R.mlp_grad(p, X, y)
        l1, g1 = R.mlp_grad_static(p, X, y)
        _close(g1, g0, f"gradient n={n}")
        assert abs(l1 - l0) <= 1e-6 * max(1.0, abs(l0)), (n, l1, l0)
        assert np.all(np.isfinite(g1)) and np.any(g1 != 0)


def test_static_grad_rejects_sizes_beyond_its_area():
    print("[case] rtod_mlp_grad_static size checks", flush=True)
    p, X, y = _inputs(3, 129)
    with pytest.raises(R.RtodError) as e:
        R.mlp_grad_static(p, X, y)
    assert e.value.status == 1 and "128" in str(e.value)
    with pytest.raises(R.RtodError):
        R.mlp_grad_static(p, X[:0], y[:0])
    # the area still serves a valid call afterwards
    l0, g0 = R.mlp_grad(p[:], X[:5], y[:5])
    l1, g1 = R.mlp_grad_static(p, X[:5], y[:5])
    _close(g1, g0, "gradient n=5 after rejections")
