#!/usr/bin/env python3
"""Writes tests/golden/golden.json: the fixtures that pin the oracle.

The reference ships no tests, fixtures or datasets (its GAU/STK/TAO data are
absent), so every vector here is produced by this build's oracle and cross-
checked at generation time by an independent method:
  * generators: SHA-256 of the seeded streams of the measured configurations;
  * walk1d-W10k-S500 slides 1..20: outlier slot lists, brute force == grid ==
    sorted-window oracle;
  * gmm3d-W10k-S500 slides 1..3: brute force == grid.
Run:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, np.float32).tobytes()).hexdigest()


def main() -> None:
    out = {"generators": {}, "streams": {}}
    for wl in (O.C1, O.C2, O.C3):
        xs = wl.data()
        out["generators"][wl.name] = {"kind": wl.kind, "seed": wl.seed, "n": wl.N,
                                      "dim": wl.dim, "sha256": sha(xs),
                                      "head": [float(v) for v in xs[:3].ravel()]}
    wl = O.C1
    xs = wl.data()
    T = 20
    brute = O.run_stream(xs, wl.W, wl.S, T, wl.R, wl.K, "brute")
    grid = O.run_stream(xs, wl.W, wl.S, T, wl.R, wl.K, "grid")
    srt = O.run_stream(xs, wl.W, wl.S, T, wl.R, wl.K, "sorted1d")
    assert all(np.array_equal(a, b) and np.array_equal(a, c) for a, b, c in zip(brute, grid, srt))
    out["streams"][wl.name] = {"W": wl.W, "S": wl.S, "R": wl.R, "K": wl.K, "slides": T,
                               "outliers": [a.tolist() for a in brute]}
    wl = O.C3
    xs = wl.data()
    T = 3
    brute = O.run_stream(xs, wl.W, wl.S, T, wl.R, wl.K, "brute")
    grid = O.run_stream(xs, wl.W, wl.S, T, wl.R, wl.K, "grid")
    assert all(np.array_equal(a, b) for a, b in zip(brute, grid))
    out["streams"][wl.name] = {"W": wl.W, "S": wl.S, "R": wl.R, "K": wl.K, "slides": T,
                               "outliers": [a.tolist() for a in brute]}
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print("wrote", os.path.join(HERE, "golden.json"))


if __name__ == "__main__":
    main()
