"""CPU oracle for RTOD's per-slide outlier detection -- TEST INFRASTRUCTURE.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package; it is the checker and the timed CPU
baseline, never the product.  The product path (``rtod_amd``) does not import
it and fails loudly when its HIP library is missing.

The C restatement lives in ``oracle/rtod_oracle.c`` (built by
``oracle/Makefile``); this module binds it with ctypes.  The synthetic
streams come from ``rtod_amd.synth`` and are re-exported here.

Parity pins (tests/test_oracle.py):
  * brute force  -- closed-form lattices, isolated points, K >= W, dense
    clusters, reference-FP64 semantics comparison;
  * sorted1d     -- equals brute force on random streams (independent method);
  * grid         -- equals brute force and sorted1d slide by slide, is
    invariant under permutation of the window, monotone in R and K;
  * generators   -- (rtod_amd.synth) pinned by the SHA-256 of their output
    (tests/golden).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "librtod_oracle.so")
_lib = None


def build() -> None:
    """Compile the oracle with its Makefile (gcc)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        L.rtod_o_d2.restype = C.c_float
        L.rtod_o_d2.argtypes = [f32p, f32p, C.c_int]
        L.rtod_o_brute.restype = C.c_int64
        L.rtod_o_brute.argtypes = [f32p, C.c_int64, C.c_int, C.c_float, C.c_int, C.c_int,
                                   C.c_int, i32p, C.c_void_p]
        L.rtod_o_sorted1d.restype = C.c_int64
        L.rtod_o_sorted1d.argtypes = [f32p, C.c_int64, C.c_float, C.c_int, i32p, C.c_void_p]
        L.rtod_o_grid_create.restype = C.c_void_p
        L.rtod_o_grid_create.argtypes = [C.c_int, C.c_int64, C.c_int64, C.c_float, C.c_int,
                                         C.c_int]
        L.rtod_o_grid_prime.restype = C.c_int
        L.rtod_o_grid_prime.argtypes = [C.c_void_p, f32p]
        L.rtod_o_grid_slide.restype = C.c_int64
        L.rtod_o_grid_slide.argtypes = [C.c_void_p, f32p, i32p]
        L.rtod_o_grid_stats.restype = None
        L.rtod_o_grid_stats.argtypes = [C.c_void_p, np.ctypeslib.ndpointer(np.int64)]
        L.rtod_o_grid_destroy.restype = None
        L.rtod_o_grid_destroy.argtypes = [C.c_void_p]
        L.rtod_o_cell_key.restype = C.c_uint64
        L.rtod_o_cell_key.argtypes = [f32p, C.c_int, C.c_float]
        L.rtod_o_mlp_forward.restype = None
        L.rtod_o_mlp_forward.argtypes = [f32p, f32p, C.c_int, C.c_int, f32p]
        L.rtod_o_mlp_grad.restype = C.c_float
        L.rtod_o_mlp_grad.argtypes = [f32p, f32p, f32p, C.c_int, C.c_int, f32p]
        L.rtod_o_mlp_loss64.restype = C.c_double
        L.rtod_o_mlp_loss64.argtypes = [np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS"),
                                        f32p, f32p, C.c_int]
        L.rtod_o_bf16_round.restype = C.c_float
        L.rtod_o_bf16_round.argtypes = [C.c_float]
        i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
        L.rtod_o_semantics_diff.restype = C.c_int64
        L.rtod_o_semantics_diff.argtypes = [f32p, C.c_int64, C.c_int, C.c_float, C.c_int, C.c_int,
                                            i64p]
        L.rtod_o_semantics_diff_stream.restype = C.c_int64
        L.rtod_o_semantics_diff_stream.argtypes = [f32p, C.c_int, C.c_int64, C.c_int64, C.c_int64,
                                                   C.c_int64, C.c_float, C.c_int, C.c_int, i64p,
                                                   i64p]
        L.rtod_o_run_stream.restype = C.c_int64
        L.rtod_o_run_stream.argtypes = [C.c_int, f32p, C.c_int, C.c_int64, C.c_int64, C.c_int64,
                                        C.c_float, C.c_int, C.c_int, i32p, i32p, C.c_int64,
                                        C.POINTER(C.c_double)]
        _lib = L
    return _lib


# --------------------------------------------------------------------------
# generators: the measured streams come from the product's synthetic-data
# library (inputs, not checked code); re-exported for the tests
# --------------------------------------------------------------------------
from rtod_amd.synth import C1, C2, C3, Workload, gen_gmm, gen_walk  # noqa: E402,F401

# --------------------------------------------------------------------------
# detection
# --------------------------------------------------------------------------
def d2(p, q) -> float:
    p = np.ascontiguousarray(p, np.float32).ravel()
    q = np.ascontiguousarray(q, np.float32).ravel()
    return float(lib().rtod_o_d2(p, q, p.size))


def brute(win: np.ndarray, R: float, K: int, semantics: int = 0, nthreads: int = 8,
          counts: bool = False):
    """Outlier slots of one window (rows in slot order)."""
    win = np.ascontiguousarray(win, np.float32)
    W, dim = win.shape
    out = np.empty(W, np.int32)
    cnt = np.empty(W, np.int32) if counts else None
    n = lib().rtod_o_brute(win.ravel(), W, dim, R, K, semantics, nthreads, out,
                           cnt.ctypes.data if counts else None)
    if n < 0:
        raise ValueError("bad arguments")
    return (out[:n].copy(), cnt) if counts else out[:n].copy()


def sorted1d(win: np.ndarray, R: float, K: int, counts: bool = False):
    win = np.ascontiguousarray(win, np.float32)
    W = win.shape[0]
    out = np.empty(W, np.int32)
    cnt = np.empty(W, np.int32) if counts else None
    n = lib().rtod_o_sorted1d(win.ravel(), W, R, K, out, cnt.ctypes.data if counts else None)
    if n < 0:
        raise ValueError("bad arguments")
    return (out[:n].copy(), cnt) if counts else out[:n].copy()


def semantics_diff(win: np.ndarray, R: float, K: int, nthreads: int = 8) -> dict:
    """Contract (FP32 d2 <= r2) vs the reference's FP64 strict rule on one
    window: points whose outlier status differs, points whose full counts
    differ, ordered pairs the two rules disagree on."""
    win = np.ascontiguousarray(win, np.float32)
    W, dim = win.shape
    o = np.zeros(3, np.int64)
    if lib().rtod_o_semantics_diff(win.ravel(), W, dim, R, K, nthreads, o) < 0:
        raise ValueError("bad arguments")
    return {"status": int(o[0]), "counts": int(o[1]), "pairs": int(o[2])}


def semantics_diff_stream(xs: np.ndarray, W: int, S: int, T: int, R: float, K: int,
                          every: int = 1, nthreads: int = 8):
    """semantics_diff over slides 1..T (every `every`-th): totals and the
    per-slide status differences (-1 for slides not checked)."""
    xs = np.ascontiguousarray(xs, np.float32)
    dim = xs.shape[1]
    if xs.shape[0] < W + T * S:
        raise ValueError("stream too short")
    per = np.zeros(max(T, 1), np.int64)
    o = np.zeros(3, np.int64)
    if lib().rtod_o_semantics_diff_stream(xs.ravel(), dim, W, S, T, every, R, K, nthreads, per,
                                          o) < 0:
        raise ValueError("bad arguments")
    return {"status": int(o[0]), "counts": int(o[1]), "pairs": int(o[2]),
            "slides_checked": int((per[:T] >= 0).sum())}, per[:T]


def cell_key(p, R: float) -> int:
    p = np.ascontiguousarray(p, np.float32).ravel()
    return int(lib().rtod_o_cell_key(p, p.size, R))


class Grid:
    """Incremental grid oracle (per-cell FIFOs), one slide at a time."""

    def __init__(self, dim: int, W: int, S: int, R: float, K: int, nthreads: int = 1):
        self.dim, self.W, self.S = dim, W, S
        self._h = lib().rtod_o_grid_create(dim, W, S, R, K, nthreads)
        if not self._h:
            raise ValueError("invalid grid configuration")
        self._buf = np.empty(W, np.int32)

    def prime(self, pts: np.ndarray) -> None:
        pts = np.ascontiguousarray(pts, np.float32).reshape(self.W, self.dim)
        lib().rtod_o_grid_prime(self._h, pts.ravel())

    def slide(self, new_pts: np.ndarray) -> np.ndarray:
        new_pts = np.ascontiguousarray(new_pts, np.float32).reshape(self.S, self.dim)
        n = lib().rtod_o_grid_slide(self._h, new_pts.ravel(), self._buf)
        return self._buf[:n].copy()

    def stats(self) -> dict:
        a = np.zeros(5, np.int64)
        lib().rtod_o_grid_stats(self._h, a)
        return {"candidates": int(a[0]), "active_cells": int(a[1]), "pair_tests": int(a[2]),
                "dense_cells": int(a[3]), "hits": int(a[4])}

    def close(self) -> None:
        if self._h:
            lib().rtod_o_grid_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


METHODS = {"grid": 0, "brute": 1, "sorted1d": 2, "brute_ref": 3}


def run_stream(xs: np.ndarray, W: int, S: int, T: int, R: float, K: int, method: str = "grid",
               nthreads: int = 8, with_time: bool = False):
    """Per-slide sorted outlier slot lists for slides 1..T of stream xs."""
    xs = np.ascontiguousarray(xs, np.float32)
    dim = xs.shape[1]
    if xs.shape[0] < W + T * S:
        raise ValueError("stream too short")
    n_out = np.zeros(max(T, 1), np.int32)
    cap = max(1, T * W)
    out = np.empty(cap, np.int32)
    us = C.c_double(0.0)
    tot = lib().rtod_o_run_stream(METHODS[method], xs.ravel(), dim, W, S, T, R, K, nthreads,
                                  n_out, out, cap, C.byref(us))
    if tot < 0:
        raise ValueError("oracle run failed (bad arguments)")
    lists, off = [], 0
    for t in range(T):
        lists.append(out[off:off + n_out[t]].copy())
        off += int(n_out[t])
    return (lists, us.value) if with_time else lists


# --------------------------------------------------------------------------
# measured configurations (synthetic, seeded)
# --------------------------------------------------------------------------


# --------------------------------------------------------------------------
# predictor MLP reference (learned scheduler)
# --------------------------------------------------------------------------
MLP_NP = 1313


def mlp_forward(params: np.ndarray, X: np.ndarray, bf16: bool = True) -> np.ndarray:
    X = np.ascontiguousarray(X, np.float32).reshape(-1, 6)
    out = np.empty(X.shape[0], np.float32)
    lib().rtod_o_mlp_forward(np.ascontiguousarray(params, np.float32), X.ravel(), X.shape[0],
                             int(bf16), out)
    return out


def mlp_grad(params: np.ndarray, X: np.ndarray, y: np.ndarray, bf16: bool = True):
    X = np.ascontiguousarray(X, np.float32).reshape(-1, 6)
    g = np.empty(MLP_NP, np.float32)
    loss = lib().rtod_o_mlp_grad(np.ascontiguousarray(params, np.float32), X.ravel(),
                                 np.ascontiguousarray(y, np.float32), X.shape[0], int(bf16), g)
    return float(loss), g


def mlp_loss64(params: np.ndarray, X: np.ndarray, y: np.ndarray) -> float:
    X = np.ascontiguousarray(X, np.float32).reshape(-1, 6)
    return float(lib().rtod_o_mlp_loss64(np.ascontiguousarray(params, np.float64), X.ravel(),
                                         np.ascontiguousarray(y, np.float32), X.shape[0]))


def bf16_round(x: float) -> float:
    return float(lib().rtod_o_bf16_round(x))
