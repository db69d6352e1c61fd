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


def test_struct_layouts_match_header():
    # rtod_config: 10 x 4 bytes; rtod_times: 6 doubles; rtod_stats: 12 int64 + 2 int32
    assert C.sizeof(R._Config) == 40
    assert C.sizeof(R._Times) == 48
    assert C.sizeof(R._Stats) == 12 * 8 + 2 * 4
