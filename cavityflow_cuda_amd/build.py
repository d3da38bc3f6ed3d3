"""Build libcavityflow_hip.so (gfx950) in-tree with hipcc.

All HIP code is compiled with -ffp-contract=off (no FMA contraction, so FP32 /
FP64 results round exactly as the reference CPU kernels do) and with packed
FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB_NAME = "libcavityflow_hip.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)

SOURCES = ["lbm_step.hip", "lbm_box.hip", "lbm_reduce.hip", "lbm_batch.hip", "lbm_probe.hip",
           "lbm_abi.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found")
    return exe


DRIVER_SRC = os.path.join(CSRC, "host", "cavity_driver.cpp")
DRIVER_PATH = os.path.join(HERE, "cavity_driver")


def build_driver(verbose: bool = False) -> str:
    """The native C++ cavity driver over the C ABI (plain host code, g++)."""
    cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), DRIVER_SRC,
           "-o", DRIVER_PATH + ".tmp", "-L", HERE, "-lcavityflow_hip",
           "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(DRIVER_PATH + ".tmp", DRIVER_PATH)
    return DRIVER_PATH


def build(verbose: bool = False) -> str:
    """Compile every source to an object and link the shared library."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    cmds, objs = [], []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        cmds.append([hipcc(), *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC,
                     *lang, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    # the translation units are independent: compile them side by side
    with ThreadPoolExecutor(max_workers=min(len(cmds), os.cpu_count() or 1)) as pool:
        for cmd in cmds:
            if verbose:
                print(" ".join(cmd), flush=True)
        for f in [pool.submit(subprocess.run, cmd, check=True) for cmd in cmds]:
            f.result()
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB_PATH)
    build_driver(verbose)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
