"""Build recipe for libcwx (HIP, gfx950) and the command-line tools.

    python -m scattersim_amd.build            # builds scattersim_amd/lib/libcwx.so and bin/

Objects go to scattersim_amd/_build/, the shared library to scattersim_amd/lib/ and the tools
(scattervolume, scatterthinvolume, scatterview, scatterviewsample,
scatterlayer) to scattersim_amd/bin/, all in-tree.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "lib")
BINDIR = os.path.join(PKG, "bin")
LIB = os.path.join(LIBDIR, "libcwx.so")

HIP_SOURCES = ["cwx_linalg.hip", "cwx_eig.hip", "cwx_field.hip", "cwx_fieldsep.hip", "cwx_solver.hip",
               "cwx_layer.hip", "cwx_mlp.hip", "cwx_capi.hip"]
HOST_SOURCES = ["cwx_tune.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 instructions stay off (hipcc's host pass warns that it does not know the feature)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build(verbose: bool = False, jobs: int = 8) -> str:
    """Compile every HIP source for gfx950, link libcwx.so and the tools.  Returns the library path."""
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(BINDIR, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, src.replace(".cpp", ".o"))
        objs.append(obj)
        cmds.append(["g++", "-O2", "-fPIC", "-std=c++17", "-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-ldl"], verbose)
    os.replace(tmp, LIB)
    host = ["g++", "-O2", "-std=c++17", f"-I{CSRC}", f"-L{LIBDIR}", "-Wl,-rpath,$ORIGIN/../lib"]
    tools = [("cli_scattervolume.cpp", "scattervolume"), ("cli_scatterview.cpp", "scatterview"),
             ("cli_scatterviewsample.cpp", "scatterviewsample"), ("cli_scatterlayer.cpp", "scatterlayer")]
    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        list(ex.map(lambda t: _run([*host, os.path.join(CSRC, t[0]), "-o", os.path.join(BINDIR, t[1]), "-lcwx"],
                                   verbose), tools))
    # scatterthinvolume is the same solver on a 2-D sample (readme.md:4,22)
    shutil.copy2(os.path.join(BINDIR, "scattervolume"), os.path.join(BINDIR, "scatterthinvolume"))
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
