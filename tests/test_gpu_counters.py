"""Counters and memory accounting through the C-ABI (the reference's
DEBUG_INFO counters and [Mem] lines, optixScan.cpp:66-93, 752-815,
1170-1187): candidates ("BVH Node / slide") and hits per slide against the
grid oracle on every slide; the context's device bytes against the arena
the build documents and against the device's own free-memory drop (the
reference measures cudaMemGetInfo deltas, optixScan.cpp:84-93)."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import rtod_amd as R

pytestmark = pytest.mark.gpu


def case(name):
    print(f"[case] {name}", flush=True)


KINDS = {"stencil": R.INDEX_CELL_STENCIL, "lbvh": R.INDEX_LBVH, "runs": R.INDEX_SLIDE_RUNS}


@pytest.mark.parametrize("wl,index", [(O.C1, "stencil"), (O.C2, "stencil"), (O.C3, "stencil"),
                                      (O.C3, "lbvh"), (O.C1, "runs"), (O.C2, "runs")],
                         ids=["C1", "C2", "C3", "C3-lbvh", "C1-runs", "C2-runs"])
def test_candidates_and_hits_every_slide(wl, index):
    """Hits = the sum over candidates of min(count, K + 1): exact on the GPU
    despite its early exits, equal to the oracle's on every slide."""
    T = 60 if index == "lbvh" else wl.T
    case(f"{wl.name} {index}: candidates and hits, {T} slides")
    xs = wl.data()
    g = O.Grid(wl.dim, wl.W, wl.S, wl.R, wl.K, nthreads=8)
    g.prime(xs[:wl.W])
    kind = KINDS[index]
    with R.Detector(wl.dim, wl.W, wl.S, wl.R, wl.K, index_kind=kind) as det:
        det.prime(xs[:wl.W])
        hits = 0
        for t in range(1, T + 1):
            new = xs[wl.W + (t - 1) * wl.S: wl.W + t * wl.S]
            got = det.slide(new)
            want = g.slide(new)
            st, so = det.stats(), g.stats()
            assert np.array_equal(got, want), t
            assert st["candidates_last"] == so["candidates"], (t, st, so)
            assert st["hits_last"] == so["hits"], (t, st["hits_last"], so["hits"])
            hits += so["hits"]
        assert st["hits_total"] == hits


@pytest.mark.parametrize("index", ["runs", "stencil"])
def test_pipelined_hits_total(index):
    wl = O.C2
    case(f"pipelined hits total, {index}")
    xs = wl.data()
    T = 60
    g = O.Grid(1, wl.W, wl.S, wl.R, wl.K)
    g.prime(xs[:wl.W])
    hits = 0
    for t in range(1, T + 1):
        g.slide(xs[wl.W + (t - 1) * wl.S: wl.W + t * wl.S])
        hits += g.stats()["hits"]
    with R.Detector(1, wl.W, wl.S, wl.R, wl.K, index_kind=KINDS[index]) as det:
        det.prime(xs[:wl.W])
        det.slides(xs[wl.W: wl.W + T * wl.S], T)
        assert det.stats()["hits_total"] == hits


def arena_bytes(dim, W, S, lbvh=False):
    """The context's device arena as the build documents it (HBM layout):
    ring, two sorted-window buffers (keys, slots, coordinates, batch
    positions), two tile-scratch runs of max(W, 2S), the batch ring of m + 3
    sorted batches with sample tables, the window sample tables, candidate /
    unit / count lists, three outlier bitmaps, the ready flags, the control
    block, the kernels' stamps and the prime copy; the LBVH's arrays."""
    m = W // S
    ne = m + 3
    tmpn = max(W, 2 * S)
    sl = 2
    while (W + (1 << sl) - 1) >> sl > 4096:
        sl += 1
    nsamp = (W + (1 << sl) - 1) >> sl
    bl = 3
    while (S + (1 << bl) - 1) >> bl > 512:
        bl += 1
    nbs = (S + (1 << bl) - 1) >> bl
    nsamp4 = (W + 3) >> 2 if sl > 2 else 0
    bm_words = (((W + 31) // 32) + 63) // 64 * 64
    b = 4 * dim * W                                           # ring
    b += 2 * (8 * W + 4 * W + 4 * dim * W + 4 * W)            # sorted windows
    b += 2 * (8 * tmpn + 4 * tmpn + 4 * dim * tmpn)           # tile scratch
    b += 8 * ne * S + 4 * ne * S + 8 * ne * nbs + 4 * dim * ne * S  # batch ring
    b += 2 * 8 * nsamp + (2 * 8 * nsamp4 if nsamp4 else 0)    # sample tables
    b += 4 * W + 8 * W + 4 * W                                # cand, units, cand_cnt
    b += 4 * 3 * bm_words                                     # bitmaps
    b += 4 * 3 * 64 * 1024                                    # ready flags
    b += 8 * 7                                                # stamps
    b += 4 * dim * W                                          # prime copy
    if lbvh:
        b += 2 * 8 * W + 2 * 4 * W + 4 * W + 4 * W + 4 * 2 * W + 4 * 12 * W + 4 * W + 4 * W
        b += 4                                                # n_whole (one int)
        b += 8 * ((W + 63) // 64)                             # candidate masks
    return b


def runs_arena_bytes(W, S):
    """The slide-run index's arena: m + 257 run-ring entries (the window and
    the slides a pipelined launch sorts and detects ahead, 128 per launch) of
    ceil(S / 8192) runs, each with its coordinates, slots, cells and cell
    starts (RL points, RL + 1 starts) and 4097 bucket starts and a 32-byte
    descriptor; the error word, the kernels' stamps and the prime copy."""
    m = W // S
    rps = (S + 8191) // 8192
    RL = (S + rps - 1) // rps
    nrun = (m + 257) * rps
    return nrun * (4 * RL + 4 * RL + 4 * 4097 + 4 * RL + 4 * (RL + 1) + 32) + 4 + 8 * 7 + 4 * W


def test_device_memory_accounting_runs():
    case("device memory, slide-run index")
    for W, S in [(100000, 5000), (100000, 10000), (10000, 500)]:
        with R.Detector(1, W, S, 0.45, 50) as det:
            st = det.stats()
            assert st["index_kind"] == R.INDEX_SLIDE_RUNS
            assert st["device_bytes"] == runs_arena_bytes(W, S), (W, S, st["device_bytes"])


@pytest.mark.parametrize("dim,W,S,lbvh", [(1, 100000, 5000, False), (3, 10000, 500, False),
                                          (3, 10000, 500, True), (2, 24576, 8192, False)])
def test_device_memory_accounting(dim, W, S, lbvh):
    case(f"device memory dim={dim} W={W} S={S} lbvh={lbvh}")
    hip = C.CDLL("libamdhip64.so")
    assert hip.hipSetDevice(0) == 0
    free0, tot = C.c_size_t(), C.c_size_t()
    kind = R.INDEX_LBVH if lbvh else R.INDEX_CELL_STENCIL
    det = R.Detector(dim, W, S, 1.0, 20, index_kind=kind)
    st = det.stats()
    assert hip.hipMemGetInfo(C.byref(free0), C.byref(tot)) == 0
    det.close()
    free1 = C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free1), C.byref(tot)) == 0
    released = free1.value - free0.value
    want = arena_bytes(dim, W, S, lbvh)
    ctl = st["device_bytes"] - want  # the control block: counters on their own lines
    assert 0 < ctl <= 4096, (st["device_bytes"], want)
    # what the device got back when the context went: at least its arena,
    # at most its allocations rounded up to the runtime's page granule
    assert st["device_bytes"] <= released <= st["device_bytes"] + 80 * (2 << 20), \
        (st["device_bytes"], released)
    assert st["host_pinned_bytes"] > 0
