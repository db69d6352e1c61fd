"""Host-side checks that need no GPU: the C-ABI library loads, exports every
symbol include/rtod.h declares, and rejects invalid configurations before
touching a device."""
import ctypes as C
import os
import re

import pytest

import rtod_amd as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    with open(os.path.join(ROOT, "include", "rtod.h")) as f:
        src = f.read()
    return sorted(set(re.findall(r"RTOD_API\s+[\w\s\*]+?\b(rtod_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = R.lib()
    names = _declared_symbols()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    # the Python mirror binds exactly the declared set
    assert sorted(n for n, _, _ in R.SIGNATURES) == names


def test_version_and_status_strings():
    assert "gfx950" in R.version()
    assert R.lib().rtod_status_str(1).decode() == "invalid argument"
    assert R.lib().rtod_status_str(2).decode() == "unsupported"


@pytest.mark.parametrize("kw,status", [
    (dict(dim=0, window=100, slide=10, R=0.5, K=5), 1),
    (dict(dim=4, window=100, slide=10, R=0.5, K=5), 1),
    (dict(dim=1, window=100, slide=30, R=0.5, K=5), 1),   # W % S != 0
    (dict(dim=1, window=100, slide=200, R=0.5, K=5), 1),  # S > W
    (dict(dim=1, window=100, slide=10, R=0.0, K=5), 1),
    (dict(dim=1, window=100, slide=10, R=float("nan"), K=5), 1),
    (dict(dim=1, window=100, slide=10, R=1e30, K=5), 1),
    (dict(dim=1, window=100, slide=10, R=0.5, K=0), 1),
    (dict(dim=1, window=0, slide=1, R=0.5, K=5), 1),
    # ablations (the reference's OPTIMIZATION / UPDATE_GAS_TYPE / --launch_ray_num)
    (dict(dim=1, window=100, slide=10, R=0.5, K=5, ablation=3), 1),
    (dict(dim=1, window=100, slide=10, R=0.5, K=5, launch_ray_num=-1,
          ablation=R.ABL_NO_FILTER), 1),
    (dict(dim=1, window=100, slide=10, R=0.5, K=5, launch_ray_num=50), 1),  # OPT0 only
    (dict(dim=1, window=100, slide=10, R=0.5, K=5, launch_ray_num=50,
          ablation=R.ABL_FILTER_ONLY), 1),
    (dict(dim=1, window=100, slide=10, R=0.5, K=5, refit=True), 2),  # stencil, default
    (dict(dim=1, window=100, slide=10, R=0.5, K=5, refit=True, ablation=R.ABL_NO_FILTER), 2),
    (dict(dim=1, window=100, slide=10, R=0.5, K=5, refit=True,
          index_kind=R.INDEX_LBVH), 2),  # the default LBVH holds the candidates
    (dict(dim=1, window=100, slide=10, R=0.5, K=5, sched_kind=R.SCHED_LEARNED,
          ablation=R.ABL_NO_FILTER), 2),
])
def test_invalid_configs_rejected_without_device(kw, status):
    with pytest.raises(R.RtodError) as e:
        R.Detector(**kw)
    assert e.value.status == status


def test_null_arguments_rejected():
    lib = R.lib()
    assert lib.rtod_create(None, C.byref(C.c_void_p())) == 1
    assert lib.rtod_prepare_cell(None) == 1
    assert lib.rtod_get_stats(None, None) == 1
    lib.rtod_destroy(None)  # no-op


def test_host_buffer_arguments_checked_without_device():
    """rtod_host_alloc refuses a size or output it cannot use before any HIP
    call; rtod_host_free takes NULL and refuses a pointer it did not hand
    out."""
    lib = R.lib()
    p = C.c_void_p()
    assert lib.rtod_host_alloc(0, C.byref(p)) == 1 and not p.value
    assert lib.rtod_host_alloc(-4, C.byref(p)) == 1
    assert lib.rtod_host_alloc(64, None) == 1
    assert lib.rtod_host_free(None) == 0
    assert lib.rtod_host_free(C.c_void_p(0x1000)) == 1
    assert b"rtod_host_alloc" in lib.rtod_last_error()


def _header_layout():
    """sizeof and field offsets of the header's structs, as a C compiler sees
    include/rtod.h (gcc on a generated program)."""
    import subprocess
    import tempfile

    structs = {"rtod_config": R._Config, "rtod_times": R._Times, "rtod_stats": R._Stats,
               "rtod_batch_tuning": R._BatchTuning}
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "rtod.h"', "int main(void){"]
    for cs, py in structs.items():
        lines.append(f'printf("{cs} size %zu\\n", sizeof({cs}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cs} {f} %zu\\n", offsetof({cs}, {f}));')
    lines.append("return 0;}")
    with tempfile.TemporaryDirectory() as td:
        src, exe = os.path.join(td, "l.c"), os.path.join(td, "l")
        open(src, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-std=c11", f"-I{os.path.join(ROOT, 'include')}", "-o", exe, src],
                       check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    got = {}
    for line in out.splitlines():
        cs, f, v = line.split()
        got[(cs, f)] = int(v)
    return structs, got


def test_struct_layouts_match_header():
    """Every ctypes mirror has the header struct's size and field offsets."""
    structs, got = _header_layout()
    for cs, py in structs.items():
        assert C.sizeof(py) == got[(cs, "size")], cs
        for f, _ in py._fields_:
            assert getattr(py, f).offset == got[(cs, f)], (cs, f)


def test_run_batch_matches_the_kernels():
    """RUN_FIRST_BATCH, RUN_BATCH, RUN_PINNED_BATCH and RUN_HOST_BATCH (a
    pipelined run's batches, used to price a launch's bytes) are the
    library's."""
    with open(os.path.join(ROOT, "rtod_amd", "csrc", "rtod_runs.h")) as f:
        src = f.read()
    m = re.search(r"constexpr int kRunBatch = (\d+);", src)
    m0 = re.search(r"constexpr int kRunFirstBatch = (\d+);", src)
    mh = re.search(r"constexpr int kRunHostBatch = (\d+);", src)
    mp = re.search(r"constexpr int kRunPinnedBatch = (\d+);", src)
    assert m and int(m.group(1)) == R.RUN_BATCH
    assert m0 and int(m0.group(1)) == R.RUN_FIRST_BATCH
    assert mh and int(mh.group(1)) == R.RUN_HOST_BATCH <= R.RUN_BATCH
    assert mp and int(mp.group(1)) == R.RUN_PINNED_BATCH <= R.RUN_BATCH


@pytest.mark.parametrize("kw,status", [
    # the learned schedule orders cell-stencil range counts only
    (dict(index_kind=R.INDEX_LBVH), 2),
    (dict(miss_threshold=-0.1), 1),
    (dict(miss_threshold=1.5), 1),
    (dict(miss_threshold=float("nan")), 1),
    (dict(window=(1 << 22) + 16, slide=16), 2),   # replay size limit
])
def test_learned_config_rejected_without_device(kw, status):
    base = dict(dim=3, window=1000, slide=100, R=1.0, K=5, sched_kind=R.SCHED_LEARNED)
    base.update(kw)
    with pytest.raises(R.RtodError) as e:
        R.Detector(**base)
    assert e.value.status == status


def test_predictor_entry_points_reject_bad_arguments_without_device():
    import numpy as np
    lib = R.lib()
    p = np.zeros(R.MLP_NPARAMS, np.float32)
    x = np.zeros(6 * 4, np.float32)
    y = np.zeros(4, np.float32)
    g = np.zeros(R.MLP_NPARAMS, np.float32)
    loss = C.c_float()
    assert lib.rtod_get_predictor(None, p, R.MLP_NPARAMS) == 1
    assert lib.rtod_set_predictor(None, p, R.MLP_NPARAMS, 0) == 1
    assert lib.rtod_mlp_grad(0, p, x, y, 0, g, C.byref(loss)) == 1     # n < 1
    assert lib.rtod_mlp_grad(0, p, x, y, 129, g, C.byref(loss)) == 1   # n > 128
    assert lib.rtod_mlp_forward(0, p, x, -1, y) == 1


@pytest.mark.parametrize("streams,status", [
    ([], 1),                                            # no streams
    ([(1, 100, 30, 0.5, 5)], 1),                        # W % S != 0
    ([(3, 1000, 100, 1.0, 5), (4, 100, 10, 0.5, 5)], 1),  # bad dim in the second
])
def test_batch_create_rejects_bad_configs_without_device(streams, status):
    with pytest.raises(R.RtodError) as e:
        R.Batch(streams)
    assert e.value.status == status


def test_batch_rejects_lbvh_and_mixed_schedules_without_device():
    import numpy as np
    lib = R.lib()
    for idx, sch in [(R.INDEX_LBVH, R.SCHED_PLAIN), (R.INDEX_LBVH, R.SCHED_LEARNED)]:
        cfgs = (R._Config * 1)(R._Config(3, 1000, 100, 1.0, 5, idx, sch, 0.25, 0, 0))
        h = C.c_void_p()
        assert lib.rtod_batch_create(cfgs, 1, 0, C.byref(h)) == 2
    # one schedule per batch
    cfgs = (R._Config * 2)(R._Config(3, 1000, 100, 1.0, 5, 0, R.SCHED_PLAIN, 0.25, 0, 0),
                           R._Config(3, 1000, 100, 1.0, 5, 0, R.SCHED_LEARNED, 0.25, 0, 0))
    assert lib.rtod_batch_create(cfgs, 2, 0, C.byref(C.c_void_p())) == 2
    assert lib.rtod_batch_create(None, 1, 0, C.byref(C.c_void_p())) == 1
    assert lib.rtod_batch_prime(None, None) == 1
    z = np.zeros(4, np.int32)
    assert lib.rtod_batch_step_device(None, z, z) == 1
    lib.rtod_batch_destroy(None)  # no-op
