"""Builds libubs_hip.so for gfx950 from csrc/*.hip with hipcc.

    python universal-beta-splatting_amd/build.py [--jobs N] [--clean]

Objects go to universal-beta-splatting_amd/_build/, the library to
universal-beta-splatting_amd/gsplat/cuda/libubs_hip.so (next to the loader).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "gsplat", "cuda", "libubs_hip.so")
PROBE_LIB = os.path.join(PKG, "gsplat", "cuda", "libubs_probe.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("UBS_OFFLOAD_ARCH", "gfx950")

CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-munsafe-fp-atomics",
    # no packed-fp32 (v_pk_*_f32) code generation
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{os.path.join(ROOT, 'include')}",
    f"-I{CSRC}",
]


def _compile(src: str, tag: str = "", extra=()) -> str:
    obj = os.path.join(BUILD, (tag or os.path.basename(src).replace(".hip", "")) + ".o")
    cmd = [HIPCC, *CFLAGS, *extra, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def _link(objs, out: str) -> None:
    tmp = out + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, out)


def build(jobs: int = 8, clean: bool = False) -> str:
    """The library, then the measurement probe library (libubs_probe.so: the
    VALU calibration kernels and the blend kernels with their pair counters,
    everything but its ubs_probe_* entry points hidden)."""
    if clean and os.path.isdir(BUILD):
        shutil.rmtree(BUILD)
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    probe = [(os.path.join(CSRC, "probe", "ubs_probe_valu.hip"), "probe_valu"),
             (os.path.join(CSRC, "ubs_raster_fwd.hip"), "probe_raster_fwd"),
             (os.path.join(CSRC, "ubs_raster_bwd.hip"), "probe_raster_bwd"),
             (os.path.join(CSRC, "ubs_capi.hip"), "probe_capi")]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        lib_jobs = [ex.submit(_compile, s) for s in srcs]
        probe_jobs = [ex.submit(_compile, s, tag, ["-DUBS_PROBE_BUILD", "-fvisibility=hidden"])
                      for s, tag in probe]
        objs = [j.result() for j in lib_jobs]
        pobjs = [j.result() for j in probe_jobs]
    _link(objs, LIB)
    _link(pobjs, PROBE_LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("--clean", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.clean))
    sys.exit(0)
