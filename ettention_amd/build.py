"""Builds libetn.so (HIP kernels + C++ host layer + C ABI) in-tree for gfx950,
and the command-line tool ettention_amd/etn linked against it.

    python -m ettention_amd.build

Objects go to ettention_amd/_build/, the library to ettention_amd/libetn.so.
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
OUT = os.path.join(PKG, "libetn.so")
CLI = os.path.join(PKG, "etn")
OBJ = os.path.join(PKG, "_build")

SOURCES = [
    "projectors.hip",
    "fused.hip",
    "fused_inst.hip",
    "unet_kernels.hip",
    "ramp_filter.hip",
    "geometry.cpp",
    "projection_set.cpp",
    "operators.cpp",
    "block_iterative.cpp",
    "batch.cpp",
    "wbp.cpp",
    "unet.cpp",
    "xml.cpp",
    "io.cpp",
    "config.cpp",
    "framework.cpp",
    "capi.cpp",
    "capi_io.cpp",
]

# fused_inst.hip is built once per (arithmetic mode, storage type): each object
# holds one sixth of the fused plane kernel variants, compiled in parallel.
FUSED_VARIANTS = [(mode, storage) for mode in ("fast", "exact") for storage in ("F32", "F16", "BF16")]


def units():
    """(source, object stem, extra flags) of every translation unit."""
    out = []
    for src in SOURCES:
        stem = os.path.splitext(src)[0]
        if src == "fused_inst.hip":
            for mode, storage in FUSED_VARIANTS:
                entry = "launch_fused_%s_%s" % (mode, storage)
                out.append((src, "fused_%s_%s" % (mode, storage.lower()),
                            ["-DETN_FUSED_FAST=%s" % ("true" if mode == "fast" else "false"),
                             "-DETN_FUSED_STORAGE=" + storage, "-DETN_FUSED_ENTRY=" + entry]))
        else:
            out.append((src, stem, []))
    return out


# -ffp-contract=off: every float expression rounds as written (the projector
# arithmetic follows the reference's OpenCL C term by term).
# -packed-fp32-ops disabled for device code (packed FP32 instructions are kept
# out of all kernels).
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "-Wall", "-Wno-unused-result",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(unit):
    src, stem, extra = unit
    obj = os.path.join(OBJ, stem + ".o")
    cmd = [hipcc()] + FLAGS + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    noise = "'-packed-fp32-ops' is not a recognized feature"
    err = "\n".join(l for l in p.stderr.splitlines() if noise not in l)
    if p.returncode != 0:
        raise RuntimeError("compiling %s failed:\n%s" % (src, err))
    if err.strip():
        sys.stderr.write(err + "\n")
    return obj


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    todo = units()
    # the slowest units first, so that they start before the pool fills up
    todo.sort(key=lambda u: (u[0] != "fused_inst.hip", u[0] != "unet_kernels.hip"))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(todo))) as ex:
        objs = list(ex.map(_compile, todo))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lz"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("linking libetn.so failed:\n" + p.stderr)
    os.replace(tmp, OUT)
    # the command-line tool: a plain C++ main over the C ABI, found next to the library
    tmp = CLI + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), os.path.join(CSRC, "etn_main.cpp"),
           "-o", tmp, "-L" + PKG, "-letn", "-Wl,-rpath,$ORIGIN"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("linking the etn tool failed:\n" + p.stderr)
    os.replace(tmp, CLI)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build()
