"""Build recipe for the HIP library (gfx950) and the C++ host driver.

    python -m 3dalm_amd.build      (or __graft_entry__.build())

Produces, in-tree:
  3dalm_amd/lib/lib3dalm_hip.so   -- the C-ABI library of include/dalm_hip.h
  3dalm_amd/lib/dalm_solver       -- the C++ case driver (setup.json/mesh.json/windturbine.json)

Every HIP translation unit is compiled with -ffp-contract=off unless it is a
precision-form unit (name ending in _mixed.hip), where contraction is allowed, and
with packed FP32 instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
LIB = os.path.join(PKG, "lib")
OBJ = os.path.join(PKG, "lib", "obj")
SO = os.path.join(LIB, "lib3dalm_hip.so")
DRIVER = os.path.join(LIB, "dalm_solver")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I" + os.path.join(ROOT, "include"), "-Wall", "-Wno-unused-function",
          "-Wno-unused-variable", "-Wno-unused-but-set-variable"]


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, src.replace(".hip", ".o"))
    contract = "-ffp-contract=fast" if src.endswith("_mixed.hip") else "-ffp-contract=off"
    cmd = [HIPCC, *COMMON, contract, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
    return obj


def build_lib(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = SO + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, SO)
    if verbose:
        print(f"built {SO}")
    return SO


def build_driver(verbose: bool = True) -> str:
    srcs = [os.path.join(HOST, f) for f in sorted(os.listdir(HOST)) if f.endswith(".cpp")] \
        if os.path.isdir(HOST) else []
    if not srcs:
        return ""
    cmd = ["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), "-o", DRIVER, *srcs,
           "-L" + LIB, "-l3dalm_hip", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver build failed:\n{r.stderr}")
    if verbose:
        print(f"built {DRIVER}")
    return DRIVER


def build(verbose: bool = True):
    build_lib(verbose)
    build_driver(verbose)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
