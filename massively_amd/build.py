"""Builds libmsv.so (hand-written gfx950 HIP kernels + the C-ABI host runtime).

Every ``csrc/*.hip`` file is compiled separately with ``hipcc`` for gfx950 and
linked into ``massively_amd/libmsv.so``. The library is built in-tree so that it
travels with the source tree.

Usage: ``python -m massively_amd.build [-j N]``
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
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libmsv.so"
ARCH = "gfx950"

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

# -packed-fp32-ops: no packed FP32 arithmetic is generated (some of its forms
# have given silently wrong results on gfx950).
COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-I", str(PKG.parent / "include"),
]


# Device headers that user-defined operators are compiled against at run time
# (hiprtc), embedded in the library as string literals under the names the
# sources include them by.
JIT_HEADERS = [("device_common.hpp", CSRC / "device_common.hpp"),
               ("scan_kernels.hpp", CSRC / "scan_kernels.hpp"),
               ("predicate.hpp", CSRC / "predicate.hpp"),
               ("select_kernels.hpp", CSRC / "select_kernels.hpp"),
               ("merge_kernels.hpp", CSRC / "merge_kernels.hpp"),
               ("../../include/msv.h", PKG.parent / "include" / "msv.h")]


def _write_jit_headers() -> None:
    names, bodies = [], []
    for name, path in JIT_HEADERS:
        text = path.read_text()
        # hiprtc supplies the HIP runtime itself (size_t included).
        text = text.replace("#include <hip/hip_runtime.h>", "").replace("#include <stddef.h>", "")
        assert ")msvjit\"" not in text
        names.append(f'"{name}"')
        bodies.append('R"msvjit(' + text + ')msvjit"')
    out = OBJ / "jit_headers.inc"
    body = ("static const char* const kJitHeaderNames[] = {" + ", ".join(names) + "};\n"
            "static const char* const kJitHeaderBodies[] = {\n" + ",\n".join(bodies) + "};\n"
            f"static const int kJitHeaderCount = {len(names)};\n")
    if not out.exists() or out.read_text() != body:
        out.write_text(body)


def _compile(src: Path) -> Path:
    obj = OBJ / (src.stem + ".o")
    cmd = [HIPCC, *COMMON, "-I", str(OBJ), "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stdout}\n{res.stderr}")
    return obj


def build(jobs: int | None = None, verbose: bool = True) -> Path:
    OBJ.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    if not sources:
        raise RuntimeError("no HIP sources found")
    _write_jit_headers()
    jobs = jobs or min(len(sources), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as pool:
        objs = list(pool.map(_compile, sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-lhiprtc", "-o",
           str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(sources)} sources", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    a = ap.parse_args()
    build(a.j)
