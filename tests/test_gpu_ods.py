"""The ods CLI (rtod_amd/ods, the counterpart of the reference's optixScan
main/parse_args/read_data, optixScan.cpp:95-212, 1128-1190): text inputs in
the reference's formats (1-D one %lf value per line, 3-D "x,y,z"), its
flags, and every slide's outlier list (--out) against the grid oracle.  The
error-path checks need no GPU (they fail before any device call)."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ODS = os.path.join(ROOT, "rtod_amd", "ods")


def run_ods(args, timeout=120):
    return subprocess.run([ODS, *map(str, args)], capture_output=True, text=True,
                          timeout=timeout)


def read_out(path):
    lists = []
    for line in open(path):
        v = [int(x) for x in line.split()]
        assert v[1] == len(v) - 2
        lists.append(np.array(v[2:], np.int32))
    return lists


def write_points(path, xs):
    with open(path, "w") as f:
        for p in xs:
            # %.9g round-trips float32 through the reference's %lf parse
            f.write(",".join(f"{float(v):.9g}" for v in p) + "\n")


def assert_same(got, want, name):
    assert len(got) == len(want), f"{name}: {len(got)} slides vs {len(want)}"
    for t, (g, w) in enumerate(zip(got, want), 1):
        assert np.array_equal(g, w), f"{name}: slide {t}"


def test_ods_rejects_unknown_flags_and_missing_files(tmp_path):
    r = run_ods(["--bogus"])
    assert r.returncode == 1 and "Unknown option '--bogus'" in r.stderr
    r = run_ods(["--file", tmp_path / "absent.txt", "--window", "10", "--slide", "5",
                 "--R", "1", "--K", "2"])
    assert r.returncode == 1 and f"Fail to open [{tmp_path / 'absent.txt'}]!" in r.stderr
    r = run_ods(["--window", "10"])
    assert r.returncode == 1 and "required" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("dim", [1, 3])
def test_ods_text_input_every_slide(dim, tmp_path):
    """read_data's formats, --n and --start_copy_pos (optixScan.cpp:95-147,
    :1136): slides of the stream after the first start_copy_pos slides."""
    print(f"[case] ods text input dim={dim}", flush=True)
    W, S, Rr, K = (2000, 100, 0.45, 10) if dim == 1 else (3000, 150, 1.9, 20)
    n, skip = W + 40 * S, 3
    xs = O.gen_walk(81, n + 500, 1) if dim == 1 else O.gen_gmm(82, n + 500, 3)
    f = tmp_path / "pts.txt"
    write_points(f, xs)
    out = tmp_path / "out.txt"
    r = run_ods(["--file", f, "--n", n, "--dim", dim, "--window", W, "--slide", S, "--R", Rr,
                 "--K", K, "--start_copy_pos", skip, "--warmup", 1, "--out", out])
    assert r.returncode == 0, r.stderr
    assert "[Time] total time for a slide:" in r.stdout and "[Mem] Device memory" in r.stdout
    assert "BVH Node / slide:" in r.stdout and "Hit / slide:" in r.stdout
    sub = xs[skip * S: n]
    T = (len(sub) - W) // S
    got = read_out(out)
    assert_same(got, O.run_stream(sub, W, S, T, Rr, K, "grid"), f"ods dim={dim}")


@pytest.mark.gpu
@pytest.mark.parametrize("extra", [[], ["--pipelined"], ["--pinned"], ["--pipelined", "--pinned"],
                                   ["--index", "lbvh"],
                                   ["--optimization", "0", "--index", "lbvh",
                                    "--update_gas_type", "0"],
                                   ["--optimization", "1", "--index", "lbvh"],
                                   ["--optimization", "0"]],
                         ids=["default", "pipelined", "pinned", "pipelined-pinned", "lbvh",
                              "opt0-lbvh-refit", "opt1-lbvh",
                              "opt0-stencil"])
def test_ods_gen_and_modes(extra, tmp_path):
    """--gen walk --seed 2 is the measured 1-D stream (the package's
    generator); every mode lists the same outliers."""
    print(f"[case] ods --gen {extra}", flush=True)
    W, S, Rr, K, n = 10000, 500, 0.45, 50, 10000 + 30 * 500
    out = tmp_path / "out.txt"
    r = run_ods(["--gen", "walk", "--seed", 2, "--n", n, "--window", W, "--slide", S, "--R", Rr,
                 "--K", K, "--warmup", 0, "--out", out, *extra])
    assert r.returncode == 0, r.stderr
    xs = O.gen_walk(2, n, 1)
    assert_same(read_out(out), O.run_stream(xs, W, S, 30, Rr, K, "grid"), f"ods {extra}")


@pytest.mark.gpu
def test_ods_launch_ray_num(tmp_path):
    """--optimization 0 --launch_ray_num L: the rays (slot i mod W) whose
    count is <= K, as the reference lists them (optixScan.cu:64-69)."""
    print("[case] ods --launch_ray_num", flush=True)
    W, S, Rr, K, n, L = 4000, 200, 0.45, 10, 4000 + 10 * 200, 9100
    out = tmp_path / "out.txt"
    r = run_ods(["--gen", "walk", "--seed", 5, "--n", n, "--window", W, "--slide", S, "--R", Rr,
                 "--K", K, "--warmup", 0, "--out", out, "--optimization", 0, "--index", "lbvh",
                 "--launch_ray_num", L])
    assert r.returncode == 0, r.stderr
    xs = O.gen_walk(5, n, 1)
    want = []
    for w in O.run_stream(xs, W, S, 10, Rr, K, "grid"):
        e = np.concatenate([w + b for b in range(0, L, W)])
        want.append(np.sort(e[e < L]).astype(np.int32))
    assert_same(read_out(out), want, "ods launch_ray_num")
