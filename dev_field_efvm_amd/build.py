"""Build the HIP library ``libefvm.so`` in-tree for gfx950 with hipcc.

The FP64 parity translation unit is compiled with ``-ffp-contract=off`` (the
reference never fuses multiply-adds); the FP32 production unit fuses a
multiply-add only inside one source expression (``-ffp-contract=on``), which the
front end decides once per expression: the four compiled copies of the row step
then round alike, and results do not depend on the strip height.  The FP32 unit also runs
with the IEEE mode bit clear, no NaN operands and no signed zeros assumed (``-fno-honor-nans
-mno-amdgpu-ieee -fno-signed-zeros``): min/max then need no operand canonicalisation, factors
2 and 1/2 ride on the output modifiers of the instructions before them, and its one NaN test
(the CFL's non-physical-cell check) is written on bit patterns.  The learned-reconstruction
unit also fuses multiply-adds per source expression only: its fused bypass stage and the
separate passes it replaces then round alike wherever they evaluate the same expression.  All
HIP code is built with packed-FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libefvm.so")
ARCH = os.environ.get("EFVM_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-fPIC", "-std=c++17", "-x", "hip", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-result", "-Wno-unused-command-line-argument"]
UNITS = [
    ("efvm_stage_f64.hip", ["-ffp-contract=off"]),
    ("efvm_stage_f32.hip", ["-ffp-contract=on", "-fgpu-flush-denormals-to-zero", "-fno-honor-nans", "-mno-amdgpu-ieee",
                            "-fno-signed-zeros"]),
    ("efvm_mlp.hip", ["-ffp-contract=on"]),
    ("efvm_stage3d.hip", ["-ffp-contract=off"]),
    ("efvm_api.cpp", ["-ffp-contract=off"]),
    ("efvm_api3d.cpp", ["-ffp-contract=off"]),
]


def hipcc() -> str:
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libefvm")


def build(verbose: bool = False, jobs: int = 0) -> str:
    """Compile every translation unit (in parallel, `jobs` at a time; default one per unit up to
    the host's cores) and link libefvm.so.  A missing source is an error."""
    os.makedirs(OBJ, exist_ok=True)
    cc = hipcc()
    cmds, objs = [], []
    for src, extra in UNITS:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        cmds.append([cc, *COMMON, *extra, "-c", path, "-o", obj])
        objs.append(obj)
    jobs = jobs or min(len(cmds), os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        for cmd, r in zip(cmds, pool.map(lambda c: subprocess.run(c, capture_output=True, text=True), cmds)):
            if verbose:
                print(" ".join(cmd), flush=True)
            errs = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
            if r.returncode != 0:
                raise RuntimeError(f"{' '.join(cmd)} failed:\n{r.stderr}")
            if verbose and errs:
                print(errs, flush=True)
    tmp = LIB + ".tmp"
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
