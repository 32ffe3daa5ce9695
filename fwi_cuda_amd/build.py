"""Compile the gfx950 library libfwi_psv.so in place (hipcc, no JIT cache).

    python -m fwi_cuda_amd.build            # library
    python -m fwi_cuda_amd.build --driver   # also the fwi_psv host driver executable
    python -m fwi_cuda_amd.build --changed  # recompile only objects older than their
                                            # source or a header of csrc/ or include/

Objects go to fwi_cuda_amd/_build/ (git-ignored); the .so lands next to this
file so it travels with the source tree.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libfwi_psv.so"
DRIVER = PKG / "fwi_psv"

HIP_SOURCES = [  # the longest compiles first
    "psv_step7_b.hip", "psv_step7_c.hip", "psv_step7.hip", "psv_stencil6.hip", "psv_v5.hip", "psv_v11.hip",
    "psv_v13.hip", "psv_kernels.hip", "psv_misc.hip", "psv_plain.hip", "psv_api.hip", "psv_noalloc.hip",
    "psv_exact.hip", "psv_fp64.hip"]
CPP_SOURCES = ["psv_preproc.cpp"]
DRIVER_SOURCES = ["host/psv_io.cpp", "host/psv_main.cpp"]

ARCH = os.environ.get("FWI_OFFLOAD_ARCH", "gfx950")
COMMON = ["-O3", "-fPIC", "-std=c++17", f"-I{ROOT / 'include'}", f"-I{CSRC}"]
# packed FP32 instructions are disabled for every HIP translation unit
HIP_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed: {' '.join(map(str, cmd))}\n{r.stdout}\n{r.stderr}")
    return r


def _stale(src: Path, obj: Path) -> bool:
    """An object is stale when it is missing or older than its source or any header."""
    if not obj.exists():
        return True
    t = obj.stat().st_mtime
    headers = list(CSRC.glob("*.h")) + list((ROOT / "include").glob("*.h")) + list((CSRC / "host").glob("*.h"))
    return any(p.stat().st_mtime > t for p in [src, *headers])


def build(driver: bool = False, verbose: bool = False, changed_only: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    cc = hipcc()
    jobs = []
    for s in HIP_SOURCES:
        o = BUILD / (Path(s).stem + ".o")
        if not changed_only or _stale(CSRC / s, o):
            jobs.append([cc, *COMMON, *HIP_FLAGS, "-c", str(CSRC / s), "-o", str(o)])
    for s in CPP_SOURCES:
        o = BUILD / (Path(s).stem + ".o")
        if not changed_only or _stale(CSRC / s, o):
            jobs.append(["g++", *COMMON, "-Wall", "-c", str(CSRC / s), "-o", str(o)])
    if jobs:
        with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
            for r in ex.map(_run, jobs):
                if verbose and r.stderr:
                    print(r.stderr, file=sys.stderr)
    objs = [str(BUILD / (Path(s).stem + ".o")) for s in HIP_SOURCES + CPP_SOURCES]
    if jobs or not LIB.exists():
        tmp = LIB.with_suffix(".so.tmp")
        _run([cc, *HIP_FLAGS, "-shared", "-fPIC", "-o", str(tmp), *objs])
        os.replace(tmp, LIB)
    if driver:
        dobjs = []
        for s in DRIVER_SOURCES:
            o = BUILD / (Path(s).stem + ".o")
            if not changed_only or _stale(CSRC / s, o):
                _run(["g++", *COMMON, "-Wall", "-c", str(CSRC / s), "-o", str(o)])
            dobjs.append(str(o))
        _run(["g++", "-o", str(DRIVER), *dobjs, f"-L{PKG}", "-lfwi_psv", f"-Wl,-rpath,$ORIGIN"])
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--driver", action="store_true")
    ap.add_argument("--changed", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(driver=a.driver, verbose=a.verbose, changed_only=a.changed))
