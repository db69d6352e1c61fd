"""Synthetic streams of the measured workloads (host/synth.c via ctypes).

Single streams:
  C1 walk1d-W10k-S500   1-D random walk, 200,000 points, W=10k, S=500, R=0.45, K=50
  C2 walk1d-W100k-S5k   1-D random walk, 1,048,572 points, W=100k, S=5k, R=0.45, K=50
  C3 gmm3d-W10k-S500    3-D mixture of 8 drifting Gaussians + 0.5% noise,
                        575,000 points, W=10k, S=500, R=1.9, K=50
  gmm3d-W100k-S5k       the same generator, 2.1M points, W=100k, S=5k, R=0.9,
                        K=50 (detection throughput-bound)
Batched (C4): `batch_workload` draws n independent streams of uneven shape --
dim ~ U{1,2,3}, W log-uniform on [2k, 100k] rounded to a multiple of S,
S = 1-5% of W, K in {20, 50} -- with R set on the first window so that the
median neighbour count is about 2K.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "librtod_synth.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"{_LIB_PATH} is missing: run rtod_amd/build.py")
        L = C.CDLL(_LIB_PATH)
        f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        L.rtod_synth_walk.restype = None
        L.rtod_synth_walk.argtypes = [C.c_uint64, C.c_int64, C.c_int, C.c_double, C.c_double,
                                      C.c_double, C.c_double, C.c_double, f32p]
        L.rtod_synth_gmm.restype = None
        L.rtod_synth_gmm.argtypes = [C.c_uint64, C.c_int64, C.c_int, C.c_int, C.c_double,
                                     C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_double, f32p]
        _lib = L
    return _lib


def gen_walk(seed: int, n: int, dim: int = 1, start: float = 10.0, sigma: float = 0.05,
             p_spike: float = 0.002, spike_lo: float = 2.0, spike_hi: float = 5.0) -> np.ndarray:
    out = np.empty(n * dim, np.float32)
    lib().rtod_synth_walk(seed, n, dim, start, sigma, p_spike, spike_lo, spike_hi, out)
    return out.reshape(n, dim)


def gen_gmm(seed: int, n: int, dim: int = 3, ncomp: int = 8, lo: float = 0.0, hi: float = 100.0,
            sig_lo: float = 0.5, sig_hi: float = 3.0, noise_frac: float = 0.005,
            drift: float = 5.0) -> np.ndarray:
    out = np.empty(n * dim, np.float32)
    lib().rtod_synth_gmm(seed, n, dim, ncomp, lo, hi, sig_lo, sig_hi, noise_frac, drift, out)
    return out.reshape(n, dim)


@dataclass(frozen=True)
class Workload:
    name: str
    dim: int
    N: int
    W: int
    S: int
    R: float
    K: int
    kind: str  # "walk" | "gmm"
    seed: int

    @property
    def T(self) -> int:
        """Number of slides: min(10000, floor((N-W)/S)) (optixScan.cpp:1036,1045)."""
        return min(10000, (self.N - self.W) // self.S)

    def data(self, n: int | None = None) -> np.ndarray:
        """The stream's first max(N, n) points (longer runs extend the seed)."""
        n = max(self.N, n or 0)
        if self.kind == "walk":
            return gen_walk(self.seed, n, self.dim)
        return gen_gmm(self.seed, n, self.dim)


C1 = Workload("walk1d-W10k-S500", 1, 200_000, 10_000, 500, 0.45, 50, "walk", 1)
C2 = Workload("walk1d-W100k-S5k", 1, 1_048_572, 100_000, 5_000, 0.45, 50, "walk", 2)
C3 = Workload("gmm3d-W10k-S500", 3, 575_000, 10_000, 500, 1.9, 50, "gmm", 3)
# a large 3-D window: ~90k candidates per slide, detection throughput-bound
C3L = Workload("gmm3d-W100k-S5k", 3, 2_100_000, 100_000, 5_000, 0.9, 50, "gmm", 11)
WORKLOADS = {w.name: w for w in (C1, C2, C3, C3L)}


@dataclass(frozen=True)
class BatchStream:
    dim: int
    W: int
    S: int
    R: float
    K: int
    kind: str
    seed: int


def calibrate_R(win: np.ndarray, K: int, queries: int = 64, seed: int = 0) -> float:
    """R such that about half of the window's points have >= 2K neighbours
    (count including the point): the median over sampled points of the
    distance to the 2K-th nearest window point (the point itself first)."""
    rng = np.random.default_rng(seed)
    q = win[rng.choice(len(win), size=min(queries, len(win)), replace=False)].astype(np.float64)
    w = win.astype(np.float64)
    kth = []
    for p in q:
        d2 = ((w - p) ** 2).sum(axis=1)
        kth.append(np.sqrt(np.partition(d2, 2 * K - 1)[2 * K - 1]))
    return float(np.float32(np.median(kth)))


def batch_workload(n_streams: int = 256, seed: int = 4, w_lo: int = 2000, w_hi: int = 100_000,
                   slides: int = 60) -> tuple[list[BatchStream], list[np.ndarray]]:
    """C4: n independent streams of uneven W, S, R, K and dim, each with
    W + slides*S points.  Returns (descriptors, point arrays)."""
    rng = np.random.default_rng(seed)
    descs, data = [], []
    for i in range(n_streams):
        dim = int(rng.integers(1, 4))
        W0 = float(np.exp(rng.uniform(np.log(w_lo), np.log(w_hi))))
        S = max(1, int(round(W0 * rng.uniform(0.01, 0.05))))
        W = max(S, int(round(W0 / S)) * S)
        K = int(rng.choice([20, 50]))
        kind = "walk" if dim == 1 or (dim == 2 and i % 2 == 0) else "gmm"
        s_seed = 1000 + i
        n = W + slides * S
        xs = gen_walk(s_seed, n, dim) if kind == "walk" else gen_gmm(s_seed, n, dim)
        R = calibrate_R(xs[:W], K, seed=i)
        descs.append(BatchStream(dim, W, S, R, K, kind, s_seed))
        data.append(xs)
    return descs, data
