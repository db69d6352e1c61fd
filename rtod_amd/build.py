"""Build the HIP library (librtod.so) and the C++ ods CLI for gfx950, in-tree.

Every HIP source is compiled with -ffp-contract=off (the neighbour predicate
must round exactly like the CPU oracle) and with packed FP32 instructions
disabled.
"""
from __future__ import annotations

import hashlib
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "librtod.so")
ODS = os.path.join(HERE, "ods")

HIP_SOURCES = ["rtod_lib.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden", "-Wall", "-Wno-unused-function", "-Wno-unused-parameter",
    f"-I{os.path.join(ROOT, 'include')}",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build failed: {' '.join(cmd[:3])} ... (exit {r.returncode})")
    # hipcc's host pass warns about the device-only target feature: harmless
    for line in (r.stderr or "").splitlines():
        if "packed-fp32-ops" in line or "not a recognized feature" in line:
            continue
        if "warning" in line or "error" in line:
            sys.stderr.write(line + "\n")


def _snapshot(dst: str) -> str:
    """Copy the HIP sources and the public header into dst (same relative
    layout) and return the SHA-256 of their contents, first 16 hex digits.
    hipcc compiles the snapshot, so a source edited while a build runs cannot
    reach one of its passes and not the other."""
    h = hashlib.sha256()
    files = [("rtod_amd/csrc", f) for f in sorted(os.listdir(CSRC))
             if f.endswith((".h", ".hip"))]
    files += [("include", f) for f in sorted(os.listdir(os.path.join(ROOT, "include")))
              if f.endswith(".h")]
    for sub, f in files:
        os.makedirs(os.path.join(dst, sub), exist_ok=True)
        with open(os.path.join(ROOT, sub, f), "rb") as fh:
            data = fh.read()
        with open(os.path.join(dst, sub, f), "wb") as fh:
            fh.write(data)
        h.update(f"{sub}/{f}\0".encode())
        h.update(data)
        h.update(b"\0")
    return h.hexdigest()[:16]


def build_lib() -> str:
    tmp = f"{LIB}.{os.getpid()}.tmp"  # one per build: concurrent builds never share it
    snap = tempfile.mkdtemp(prefix="rtod_src_")
    try:
        digest = _snapshot(snap)
        srcs = [os.path.join(snap, "rtod_amd", "csrc", s) for s in HIP_SOURCES]
        flags = [f if not f.startswith("-I") else f"-I{os.path.join(snap, 'include')}"
                 for f in HIP_FLAGS]
        _run([HIPCC, *flags, f'-DRTOD_SRC_HASH="{digest}"', "-shared", "-o", tmp, *srcs])
    finally:
        shutil.rmtree(snap, ignore_errors=True)
    os.replace(tmp, LIB)
    return LIB


def build_ods() -> str:
    src = os.path.join(HERE, "host", "ods.cpp")
    if not os.path.exists(src):
        return ""
    tmp = f"{ODS}.{os.getpid()}.tmp"
    _run(["g++", "-O2", "-std=c++17", f"-I{os.path.join(ROOT, 'include')}", "-o", tmp, src,
          f"-L{HERE}", "-lrtod", "-lrtod_synth", f"-Wl,-rpath,$ORIGIN", "-lpthread"])
    os.replace(tmp, ODS)
    return ODS


SYNTH = os.path.join(HERE, "librtod_synth.so")


def build_synth() -> str:
    """Synthetic-stream generators (plain C, gcc)."""
    src = os.path.join(HERE, "host", "synth.c")
    tmp = f"{SYNTH}.{os.getpid()}.tmp"
    _run(["gcc", "-O2", "-std=c11", "-ffp-contract=off", "-fPIC", "-shared",
          "-fvisibility=hidden", "-o", tmp, src, "-lm"])
    os.replace(tmp, SYNTH)
    return SYNTH


def build_all() -> None:
    build_lib()
    build_synth()
    build_ods()


if __name__ == "__main__":
    build_all()
    print(LIB)
