"""Build recipe for the in-tree native libraries.

``python -m ddhdg_amd.build`` (or ``__graft_entry__.build()``) compiles
  ddhdg_amd/libddhdg_hip.so   the gfx950 kernels + C ABI (include/ddhdg_hip.h)
with hipcc.  Every HIP translation unit is built for gfx950 only, with packed
FP32 instructions disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libddhdg_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I" + os.path.join(ROOT, "include"), "-Wall", "-Wno-unused-function"]
SOURCES = ["hdg_tables.cpp", "hdg_topology.cpp", "hdg_nd_plan.cpp", "hdg_kernels.hip",
           "hdg_fast2d.hip", "hdg_fast2dp.hip", "hdg_fast3d.hip", "hdg_capi.hip", "hdg_trace_solver.hip",
           "hdg_local.hip", "hdg_gmres.hip"]
LIBS = ["-L/opt/rocm/lib", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build failed: {' '.join(cmd)}")
    return r


def build(verbose: bool = True) -> str:
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    cmds, objs = [], []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmds.append([HIPCC, *COMMON, *lang, "-c", path, "-o", obj])
        objs.append(obj)
    # translation units compile independently: a few at a time
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "0")) or (os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, *LIBS])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
