"""Build the native libraries in-tree (``python -m mkpp_amd.build``).

Outputs (git-ignored, shipped with the source tree):
  mkpp_amd/lib/libmkpp_amd.so        every bundled mechanism + registry + C ABI
  mkpp_amd/lib/libmkpp_<mech>.so     reference-named drop-in C ABI per mechanism
HIP code is compiled for gfx950 only.
"""

from __future__ import annotations

import glob
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
OBJ = os.path.join(HERE, "lib", "obj")
ROOT = os.path.dirname(HERE)

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-Wno-pass-failed",
]


def sources() -> list[str]:
    srcs = sorted(glob.glob(os.path.join(CSRC, "gen", "mech_*.hip")))
    srcs += sorted(glob.glob(os.path.join(CSRC, "gen", "batch_*.hip")))
    srcs += sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    srcs.append(os.path.join(CSRC, "api.cpp"))
    return srcs


def bundled_mechanisms() -> list[str]:
    return [os.path.basename(p)[len("mech_"):-len(".hip")] for p in sorted(glob.glob(os.path.join(CSRC, "gen", "mech_*.hip")))]


# The persistent batch kernel loops over work items.  Machine-level loop
# invariant code motion hoists the math library's polynomial constants out of
# that loop, which pushes the kernel past 256 VGPRs and spills them to scratch
# once per item; without it the batch kernel has no VGPR spills.  The
# one-cell-per-wave kernels have no such loop and keep the pass.
BATCH_FLAGS = ["-mllvm", "-disable-machine-licm"]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    extra = BATCH_FLAGS if os.path.basename(src).startswith("batch_") else []
    cmd = [HIPCC, *HIP_FLAGS, *extra, "-x", "hip", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose: bool = True, jobs: int | None = None) -> None:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    out = os.path.join(LIB, "libmkpp_amd.so")
    tmp = out + ".tmp"
    r = subprocess.run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, out)
    for mech in bundled_mechanisms():
        shim = os.path.join(LIB, f"libmkpp_{mech}.so")
        cmd = ["g++", "-O2", "-shared", "-fPIC", f'-DMKPP_MECH="{mech}"', os.path.join(CSRC, "capi_shim.cpp"),
               "-L" + LIB, "-lmkpp_amd", "-Wl,-rpath,$ORIGIN", "-o", shim + ".tmp"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"shim build failed for {mech}:\n{r.stderr[-4000:]}")
        os.replace(shim + ".tmp", shim)
    shutil.rmtree(OBJ, ignore_errors=True)
    if verbose:
        print(f"built {out} ({len(srcs)} units) and drop-ins for {', '.join(bundled_mechanisms())}", file=sys.stderr)


if __name__ == "__main__":
    build()
