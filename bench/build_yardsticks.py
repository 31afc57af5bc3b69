"""Builds the benchmark's measuring sticks (never linked into libmsv):

* ``bench/proxy/libmsv_refproxy.so``: a HIP rendering of the reference's radix
  schedule for u32 sort_by_key, timed as a stand-in for its CubeCL-HIP build;
* ``bench/vendor/librocprim_yardstick.so``: rocPRIM's radix sort, segmented
  radix sort and inclusive scan, timed as the vendor library's figures.

Usage: ``python bench/build_yardsticks.py``
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-shared"]
TARGETS = [(HERE / "proxy" / "reference_schedule.hip", HERE / "proxy" / "libmsv_refproxy.so"),
           (HERE / "vendor" / "rocprim_yardstick.hip", HERE / "vendor" / "librocprim_yardstick.so")]


def build(verbose: bool = True) -> None:
    for src, lib in TARGETS:
        tmp = lib.with_suffix(".so.tmp")
        res = subprocess.run([HIPCC, *FLAGS, str(src), "-o", str(tmp)], capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stdout}\n{res.stderr}")
        os.replace(tmp, lib)
        if verbose:
            print(f"built {lib}", file=sys.stderr)


if __name__ == "__main__":
    build()
