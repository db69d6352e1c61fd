"""rtod_amd -- MI355X-native RTOD per-slide outlier detection.

Python host mirror of the reference's ods slide loop over the C-ABI of
``include/rtod.h`` (library ``rtod_amd/librtod.so``, built by
``rtod_amd/build.py``).  The GPU path is the only path: there is no CPU
fallback, and loading fails loudly when the HIP library is missing.

Reference mapping (src/ods/optixScan.cpp):
  Detector(...)              <- parse_args / initialize_params / initialize_cell
                                (:149-212, :674-724, :960-965)
  Detector.prime(pts)        <- initial window load of detect_outlier (:1020-1033)
  Detector.slide(new_pts)    <- one iteration of the slide loop (:1045-1110)
  Detector.copy_new_points / prepare_cell / build_index / detect_outlier /
  copy_outlier               <- the loop body's calls (:1051-1101)
  Detector.times             <- Timer buckets (timer.h:16-24, :102-117)
"""
from __future__ import annotations

import ctypes as C
import os
import weakref
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librtod.so")

INDEX_AUTO, INDEX_CELL_STENCIL, INDEX_LBVH, INDEX_SLIDE_RUNS = 0, 1, 2, 3
SCHED_PLAIN, SCHED_LEARNED = 0, 1
ABL_NONE, ABL_FILTER_ONLY, ABL_NO_FILTER = 0, 1, 2
FLAG_PHASE_TIMES = 0x1
FLAG_LEARN_ALWAYS = 0x2
# a pipelined run's batches on the slide-run index: the first RUN_FIRST_BATCH
# slides, then RUN_BATCH each from HBM, RUN_PINNED_BATCH from a pinned
# buffer, RUN_HOST_BATCH from pageable host memory; a launch detects one
# batch and sorts the next (kRunFirstBatch, kRunBatch, kRunPinnedBatch and
# kRunHostBatch of rtod_amd/csrc/rtod_runs.h)
RUN_FIRST_BATCH = 32
RUN_BATCH = 128
RUN_PINNED_BATCH = 64
RUN_HOST_BATCH = 32

STATUS = {0: "ok", 1: "invalid argument", 2: "unsupported", 3: "call out of order",
          4: "HIP runtime error", 5: "out of memory", 6: "internal error"}


class RtodError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class _Config(C.Structure):
    _fields_ = [("dim", C.c_int32), ("window", C.c_int32), ("slide", C.c_int32),
                ("R", C.c_float), ("K", C.c_int32), ("index_kind", C.c_int32),
                ("sched_kind", C.c_int32), ("miss_threshold", C.c_float),
                ("flags", C.c_int32), ("device", C.c_int32), ("ablation", C.c_int32),
                ("refit", C.c_int32), ("launch_ray_num", C.c_int32)]


class _Times(C.Structure):
    _fields_ = [("copy_h2d", C.c_double), ("prepare_cell", C.c_double),
                ("build_index", C.c_double), ("detect_outlier", C.c_double),
                ("copy_d2h", C.c_double), ("total", C.c_double), ("launch", C.c_double)]


class _Stats(C.Structure):
    _fields_ = [("slides", C.c_int64), ("candidates_last", C.c_int64),
                ("candidates_total", C.c_int64), ("pair_tests_last", C.c_int64),
                ("pair_tests_total", C.c_int64), ("outliers_last", C.c_int64),
                ("outliers_total", C.c_int64), ("device_bytes", C.c_int64),
                ("host_pinned_bytes", C.c_int64), ("predictor_misses", C.c_int64),
                ("predictor_points", C.c_int64), ("fallbacks", C.c_int64),
                ("train_steps", C.c_int64), ("train_loss_last", C.c_double),
                ("index_kind", C.c_int32), ("sched_kind", C.c_int32),
                ("hits_last", C.c_int64), ("hits_total", C.c_int64),
                ("plain_kernel_slides", C.c_int64)]

MLP_NPARAMS = 1313


class _BatchTuning(C.Structure):
    _fields_ = [("detect_points_per_wg", C.c_int32), ("learned_dims", C.c_int32),
                ("cell_scale", C.c_float * 3), ("us_per_step", C.c_double)]


_PP = C.POINTER(C.c_void_p)


# (name, restype, argtypes) of every symbol include/rtod.h declares
_F32P = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
SIGNATURES = [
    ("rtod_create", C.c_int, [C.POINTER(_Config), C.POINTER(C.c_void_p)]),
    ("rtod_prime", C.c_int, [C.c_void_p, _F32P, C.c_int64]),
    ("rtod_slide", C.c_int, [C.c_void_p, _F32P, _I32P, C.POINTER(C.c_int32),
                             C.POINTER(_Times)]),
    ("rtod_slide_device", C.c_int, [C.c_void_p, C.c_void_p, _I32P, C.POINTER(C.c_int32),
                                    C.POINTER(_Times)]),
    ("rtod_slides_device", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, _I32P, _I32P,
                                     C.c_int64]),
    ("rtod_slides", C.c_int, [C.c_void_p, _F32P, C.c_int32, _I32P, _I32P, C.c_int64]),
    ("rtod_host_alloc", C.c_int, [C.c_int64, C.POINTER(C.c_void_p)]),
    ("rtod_host_free", C.c_int, [C.c_void_p]),
    ("rtod_copy_new_points", C.c_int, [C.c_void_p, _F32P]),
    ("rtod_prepare_cell", C.c_int, [C.c_void_p]),
    ("rtod_build_index", C.c_int, [C.c_void_p]),
    ("rtod_detect_outlier", C.c_int, [C.c_void_p]),
    ("rtod_copy_outlier", C.c_int, [C.c_void_p, _I32P, C.POINTER(C.c_int32)]),
    ("rtod_get_stats", C.c_int, [C.c_void_p, C.POINTER(_Stats)]),
    ("rtod_kernel_times", C.c_int, [C.c_void_p, C.POINTER(C.c_char_p),
                                    C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                    C.POINTER(C.c_int32)]),
    ("rtod_reset_kernel_times", C.c_int, [C.c_void_p]),
    ("rtod_stage_times", C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.c_int32]),
    ("rtod_get_predictor", C.c_int, [C.c_void_p, _F32P, C.c_int32]),
    ("rtod_set_predictor", C.c_int, [C.c_void_p, _F32P, C.c_int32, C.c_int32]),
    ("rtod_mlp_forward", C.c_int, [C.c_int32, _F32P, _F32P, C.c_int32, _F32P]),
    ("rtod_mlp_grad", C.c_int, [C.c_int32, _F32P, _F32P, _F32P, C.c_int32, _F32P,
                                C.POINTER(C.c_float)]),
    ("rtod_mlp_grad_static", C.c_int, [C.c_int32, _F32P, _F32P, _F32P, C.c_int32, _F32P,
                                       C.POINTER(C.c_float)]),
    ("rtod_batch_create", C.c_int, [C.POINTER(_Config), C.c_int32, C.c_int32,
                                    C.POINTER(C.c_void_p)]),
    ("rtod_batch_prime", C.c_int, [C.c_void_p, _PP]),
    ("rtod_batch_step", C.c_int, [C.c_void_p, _PP, _I32P, _I32P]),
    ("rtod_batch_bind_device", C.c_int, [C.c_void_p, _PP, C.POINTER(C.c_int64)]),
    ("rtod_batch_prime_device", C.c_int, [C.c_void_p]),
    ("rtod_batch_step_device", C.c_int, [C.c_void_p, _I32P, _I32P]),
    ("rtod_batch_autotune", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(_BatchTuning)]),
    ("rtod_batch_get_tuning", C.c_int, [C.c_void_p, C.POINTER(_BatchTuning)]),
    ("rtod_batch_set_tuning", C.c_int, [C.c_void_p, C.POINTER(_BatchTuning)]),
    ("rtod_batch_stream_stats", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(_Stats)]),
    ("rtod_batch_destroy", None, [C.c_void_p]),
    ("rtod_destroy", None, [C.c_void_p]),
    ("rtod_status_str", C.c_char_p, [C.c_int]),
    ("rtod_last_error", C.c_char_p, []),
    ("rtod_version", C.c_char_p, []),
]

_lib = None


def lib():
    """Load librtod.so (raises if it was not built -- there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                               f"`python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(status: int) -> None:
    if status != 0:
        raise RtodError(status, lib().rtod_last_error().decode())


@dataclass
class Times:
    """Per-slide breakdown in ms, reference Timer bucket names."""
    copy_h2d: float
    prepare_cell: float
    build_index: float
    detect_outlier: float
    copy_d2h: float
    total: float
    launch: float  # total minus the five buckets (launch, dispatch, polling)


class Detector:
    """One sliding-window stream on one GPU (the reference's ScanState)."""

    def __init__(self, dim: int, window: int, slide: int, R: float, K: int,
                 index_kind: int = INDEX_AUTO, sched_kind: int = SCHED_PLAIN,
                 miss_threshold: float = 0.25, flags: int = 0, device: int = 0,
                 ablation: int = ABL_NONE, refit: bool = False, launch_ray_num: int = 0):
        self.dim, self.W, self.S, self.R, self.K = dim, window, slide, R, K
        cfg = _Config(dim, window, slide, R, K, index_kind, sched_kind, miss_threshold, flags,
                      device, ablation, 1 if refit else 0, launch_ray_num)
        h = C.c_void_p()
        _check(lib().rtod_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self.cap = max(window, launch_ray_num)  # entries a slide's list may hold
        self._out = np.empty(self.cap, np.int32)
        self._times = _Times()

    # -- lifecycle ---------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            lib().rtod_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- stream ------------------------------------------------------------
    def prime(self, pts: np.ndarray) -> None:
        pts = np.ascontiguousarray(pts, np.float32).reshape(-1)
        if pts.size != self.W * self.dim:
            raise ValueError("prime needs exactly W points")
        _check(lib().rtod_prime(self._h, pts, self.W))

    def slide(self, new_pts: np.ndarray) -> np.ndarray:
        """Outlier ring slots (ascending) of the window after this slide."""
        new_pts = np.ascontiguousarray(new_pts, np.float32).reshape(-1)
        if new_pts.size != self.S * self.dim:
            raise ValueError("slide needs exactly S points")
        n = C.c_int32(0)
        _check(lib().rtod_slide(self._h, new_pts, self._out, C.byref(n), C.byref(self._times)))
        return self._out[:n.value].copy()

    def slide_device(self, dev_ptr: int) -> np.ndarray:
        """Like slide(), the new points given as a device (HBM) pointer."""
        n = C.c_int32(0)
        _check(lib().rtod_slide_device(self._h, C.c_void_p(dev_ptr), self._out, C.byref(n),
                                       C.byref(self._times)))
        return self._out[:n.value].copy()

    def _out_arrays(self, n_slides: int):
        return np.zeros(n_slides + 1, np.int32), np.empty(max(1, n_slides * self.cap), np.int32)

    @staticmethod
    def split(offs: np.ndarray, slots: np.ndarray) -> list[np.ndarray]:
        """Per-slide lists from the (offsets, slots) arrays of a pipelined run."""
        return [slots[offs[i]:offs[i + 1]].copy() for i in range(len(offs) - 1)]

    def slides_device_raw(self, dev_ptr: int, n_slides: int, out=None):
        """rtod_slides_device itself: (offsets [n+1], slots); slide i's
        outliers are slots[offsets[i]:offsets[i+1]].  out: reusable arrays
        from an earlier call with as many slides."""
        offs, slots = out if out is not None else self._out_arrays(n_slides)
        _check(lib().rtod_slides_device(self._h, C.c_void_p(dev_ptr), n_slides, offs, slots,
                                        slots.size))
        return offs, slots

    def slides_raw(self, pts: np.ndarray, n_slides: int, out=None):
        """rtod_slides itself (pts contiguous float32): (offsets, slots) as
        slides_device_raw."""
        if pts.dtype != np.float32 or not pts.flags.c_contiguous:
            raise ValueError("pts must be C-contiguous float32")
        if pts.size < n_slides * self.S * self.dim:
            raise ValueError("pts holds fewer than n_slides slides")
        offs, slots = out if out is not None else self._out_arrays(n_slides)
        _check(lib().rtod_slides(self._h, pts.reshape(-1), n_slides, offs, slots, slots.size))
        return offs, slots

    def slides_device(self, dev_ptr: int, n_slides: int) -> list[np.ndarray]:
        """n_slides consecutive slides from HBM (slide i at dev_ptr + i*S*dim
        floats), pipelined; the same lists as n_slides slide_device() calls."""
        return self.split(*self.slides_device_raw(dev_ptr, n_slides))

    def slides(self, pts: np.ndarray, n_slides: int) -> list[np.ndarray]:
        """n_slides consecutive slides from host memory (slide i = rows
        i*S .. (i+1)*S of pts), pipelined, each staged into pinned memory just
        before its ingest; the same lists as n_slides slide() calls."""
        pts = np.ascontiguousarray(pts, np.float32)
        return self.split(*self.slides_raw(pts, n_slides))

    @property
    def times(self) -> Times:
        t = self._times
        return Times(t.copy_h2d, t.prepare_cell, t.build_index, t.detect_outlier, t.copy_d2h,
                     t.total, t.launch)

    # -- phase-level API (reference loop body) ------------------------------
    def copy_new_points(self, new_pts: np.ndarray) -> None:
        new_pts = np.ascontiguousarray(new_pts, np.float32).reshape(-1)
        if new_pts.size != self.S * self.dim:
            raise ValueError("slide needs exactly S points")
        _check(lib().rtod_copy_new_points(self._h, new_pts))

    def prepare_cell(self) -> None:
        _check(lib().rtod_prepare_cell(self._h))

    def build_index(self) -> None:
        _check(lib().rtod_build_index(self._h))

    def detect_outlier(self) -> None:
        _check(lib().rtod_detect_outlier(self._h))

    def copy_outlier(self) -> np.ndarray:
        n = C.c_int32(0)
        _check(lib().rtod_copy_outlier(self._h, self._out, C.byref(n)))
        return self._out[:n.value].copy()

    # -- counters ----------------------------------------------------------
    def stats(self) -> dict:
        s = _Stats()
        _check(lib().rtod_get_stats(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in _Stats._fields_}

    def kernel_times(self) -> dict:
        cap = 16
        names = (C.c_char_p * cap)()
        avg = (C.c_double * cap)()
        cnt = (C.c_int64 * cap)()
        n = C.c_int32(cap)
        _check(lib().rtod_kernel_times(self._h, names, avg, cnt, C.byref(n)))
        return {names[i].decode(): (avg[i], cnt[i]) for i in range(n.value)}

    def get_predictor(self) -> np.ndarray:
        """Current predictor parameters (learned schedule only)."""
        out = np.empty(MLP_NPARAMS, np.float32)
        _check(lib().rtod_get_predictor(self._h, out, MLP_NPARAMS))
        return out

    def set_predictor(self, params: np.ndarray, frozen: bool = False) -> None:
        """Replace the predictor (optimiser state cleared); frozen stops training."""
        p = np.ascontiguousarray(params, np.float32).ravel()
        _check(lib().rtod_set_predictor(self._h, p, p.size, 1 if frozen else 0))

    def reset_kernel_times(self) -> None:
        _check(lib().rtod_reset_kernel_times(self._h))

    STAGES = ("ingest", "window_merge", "classify", "lbvh_build", "detect", "epilogue")

    def stage_times(self) -> dict:
        """Kernel durations (ms, device clock) of the last slide run one per call."""
        out = (C.c_double * 6)()
        _check(lib().rtod_stage_times(self._h, out, 6))
        return dict(zip(self.STAGES, list(out)))


def run_stream(xs: np.ndarray, W: int, S: int, T: int, R: float, K: int, **kw):
    """Per-slide outlier slot lists for slides 1..T of stream xs ([N][dim])."""
    xs = np.ascontiguousarray(xs, np.float32)
    dim = xs.shape[1]
    with Detector(dim, W, S, R, K, **kw) as det:
        det.prime(xs[:W])
        return [det.slide(xs[W + (t - 1) * S: W + t * S]) for t in range(1, T + 1)]


def pinned_empty(n_points: int, dim: int = 1) -> np.ndarray:
    """A float32 [n_points][dim] array in pinned host memory (rtod_host_alloc,
    released by rtod_host_free once no view of it is left).  Slides that
    Detector.slide / slides / slides_raw take from it are read by the device
    in place over PCIe, with no copy into the context's staging."""
    n = int(n_points) * int(dim)
    p = C.c_void_p()
    _check(lib().rtod_host_alloc(n * 4, C.byref(p)))
    buf = (C.c_float * n).from_address(p.value)
    weakref.finalize(buf, lib().rtod_host_free, C.c_void_p(p.value))
    return np.ctypeslib.as_array(buf).reshape(int(n_points), int(dim))


class DeviceStream:
    """A whole stream uploaded once to HBM (hipMalloc'd, freed on close), so
    slides can be fed with Detector.slide_device without a per-slide copy."""

    def __init__(self, xs: np.ndarray, device: int = 0):
        xs = np.ascontiguousarray(xs, np.float32)
        self.rows, self.dim = xs.shape
        self._hip = C.CDLL("libamdhip64.so")
        self._hip.hipSetDevice(device)
        self.ptr = C.c_void_p()
        if self._hip.hipMalloc(C.byref(self.ptr), C.c_size_t(xs.nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
        if self._hip.hipMemcpy(self.ptr, xs.ctypes.data_as(C.c_void_p), C.c_size_t(xs.nbytes),
                               1) != 0:
            raise RuntimeError("hipMemcpy failed")

    def at(self, row: int) -> int:
        return self.ptr.value + row * self.dim * 4

    def close(self) -> None:
        if self.ptr and self.ptr.value:
            self._hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _tuning_dict(t: _BatchTuning) -> dict:
    return {"detect_points_per_wg": t.detect_points_per_wg, "learned_dims": t.learned_dims,
            "cell_scale": [float(x) for x in t.cell_scale], "us_per_step": t.us_per_step}


class Batch:
    """Many independent streams advanced together: each phase of a slide is
    one launch over all streams, the step one hipGraph replay (rtod_batch_*).
    streams: list of (dim, W, S, R, K); sched_kind LEARNED shares one
    online-trained predictor among all streams."""

    def __init__(self, streams, device: int = 0, sched_kind: int = SCHED_PLAIN,
                 miss_threshold: float = 0.25):
        n = len(streams)
        cfgs = (_Config * n)()
        for i, (dim, W, S, R, K) in enumerate(streams):
            cfgs[i] = _Config(dim, W, S, R, K, INDEX_AUTO, sched_kind, miss_threshold, 0,
                              device)
        self.streams = [tuple(x) for x in streams]
        self.n = n
        h = C.c_void_p()
        _check(lib().rtod_batch_create(cfgs, n, device, C.byref(h)))
        self._h = h
        self._offs = np.zeros(n + 1, np.int32)
        self._slots = np.zeros(sum(s[1] for s in streams), np.int32)
        self._keep = []

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            lib().rtod_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _ptrs(arrays):
        arrs = [np.ascontiguousarray(a, np.float32) for a in arrays]
        return arrs, (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])

    def _split(self):
        o = self._offs
        return [self._slots[o[i]:o[i + 1]].copy() for i in range(self.n)]

    def prime(self, windows) -> None:
        arrs, pp = self._ptrs(windows)
        for a, (dim, W, S, R, K) in zip(arrs, self.streams):
            assert a.size == W * dim
        _check(lib().rtod_batch_prime(self._h, C.cast(pp, _PP)))

    def step(self, new_pts):
        arrs, pp = self._ptrs(new_pts)
        for a, (dim, W, S, R, K) in zip(arrs, self.streams):
            assert a.size == S * dim
        _check(lib().rtod_batch_step(self._h, C.cast(pp, _PP), self._offs, self._slots))
        return self._split()

    def bind_device(self, dev_ptrs, n_points) -> None:
        """dev_ptrs: device addresses of each stream's [n][dim] float32 points."""
        pp = (C.c_void_p * self.n)(*dev_ptrs)
        npts = (C.c_int64 * self.n)(*n_points)
        self._keep = [pp, npts]
        _check(lib().rtod_batch_bind_device(self._h, C.cast(pp, _PP), npts))

    def prime_device(self) -> None:
        _check(lib().rtod_batch_prime_device(self._h))

    def step_device(self):
        _check(lib().rtod_batch_step_device(self._h, self._offs, self._slots))
        return self._split()

    def autotune(self, slides: int) -> dict:
        t = _BatchTuning()
        _check(lib().rtod_batch_autotune(self._h, slides, C.byref(t)))
        return _tuning_dict(t)

    def get_tuning(self) -> dict:
        t = _BatchTuning()
        _check(lib().rtod_batch_get_tuning(self._h, C.byref(t)))
        return _tuning_dict(t)

    def set_tuning(self, detect_points_per_wg: int, cell_scale=(1.0, 1.0, 1.0),
                   learned_dims: int = 6) -> None:
        """learned_dims (learned batches): bit d-1 set, the d-D streams run
        the learned kernels (bits 1 and 2 only)."""
        t = _BatchTuning(detect_points_per_wg, learned_dims, (C.c_float * 3)(*cell_scale), 0.0)
        _check(lib().rtod_batch_set_tuning(self._h, C.byref(t)))

    def stream_stats(self, i: int) -> dict:
        st = _Stats()
        _check(lib().rtod_batch_stream_stats(self._h, i, C.byref(st)))
        return {f: getattr(st, f) for f, _ in _Stats._fields_}


def mlp_forward(params: np.ndarray, X: np.ndarray, device: int = 0) -> np.ndarray:
    """The detector's predictor forward pass (matrix cores) on given features."""
    X = np.ascontiguousarray(X, np.float32).reshape(-1, 6)
    out = np.empty(X.shape[0], np.float32)
    _check(lib().rtod_mlp_forward(device, np.ascontiguousarray(params, np.float32).ravel(),
                                  X.ravel(), X.shape[0], out))
    return out


def mlp_grad(params: np.ndarray, X: np.ndarray, y: np.ndarray, device: int = 0):
    """Loss and gradient by the detector's training kernels (n <= 128)."""
    X = np.ascontiguousarray(X, np.float32).reshape(-1, 6)
    g = np.empty(MLP_NPARAMS, np.float32)
    loss = C.c_float(0.0)
    _check(lib().rtod_mlp_grad(device, np.ascontiguousarray(params, np.float32).ravel(),
                               X.ravel(), np.ascontiguousarray(y, np.float32).ravel(),
                               X.shape[0], g, C.byref(loss)))
    return float(loss.value), g


def mlp_grad_static(params: np.ndarray, X: np.ndarray, y: np.ndarray, device: int = 0):
    """mlp_grad through rtod_mlp_grad_static (the library's fixed scratch
    area, no allocation; n <= 128)."""
    X = np.ascontiguousarray(X, np.float32).reshape(-1, 6)
    g = np.empty(MLP_NPARAMS, np.float32)
    loss = C.c_float(0.0)
    _check(lib().rtod_mlp_grad_static(device, np.ascontiguousarray(params, np.float32).ravel(),
                                      X.ravel(), np.ascontiguousarray(y, np.float32).ravel(),
                                      X.shape[0], g, C.byref(loss)))
    return float(loss.value), g


def version() -> str:
    return lib().rtod_version().decode()
