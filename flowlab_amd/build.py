"""Builds libflowhip.so (gfx950) in-tree from flowlab_amd/csrc/*.hip.

Each translation unit is compiled by hipcc in parallel, then linked into
flowlab_amd/libflowhip.so.  All HIP code is compiled with packed-FP32
instructions disabled (-target-feature -packed-fp32-ops).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libflowhip.so")
ARCH = os.environ.get("FLOWHIP_ARCH", "gfx950")

SOURCES = ["fh_advect.hip", "fh_project.hip", "fh_mg.hip", "fh_mgfused.hip", "fh_engine.hip",
           "fh_unet_kernels.hip", "fh_unet.hip", "fh_batch.hip", "fh_noalloc.hip"]

HIPCC_FLAGS = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: install ROCm or set HIPCC")


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    addon = build_node_addon(verbose)
    if verbose and addon:
        print(f"built {addon}")
    return LIB


NODE_DIR = os.path.join(HERE, "node")
ADDON = os.path.join(NODE_DIR, "flowhip.node")
NODE_INCLUDE = os.environ.get("NODE_INCLUDE", "/usr/include/node")


def build_node_addon(verbose: bool = False):
    """The N-API addon (flowlab_amd/node/flowhip.node) over libflowhip.so,
    built with the host C compiler against node_api.h; skipped when Node's
    headers are absent."""
    if not os.path.exists(os.path.join(NODE_INCLUDE, "node_api.h")):
        if verbose:
            print(f"node_api.h not found under {NODE_INCLUDE}: N-API addon not built")
        return None
    tmp = ADDON + ".tmp"
    cmd = ["gcc", "-O2", "-shared", "-fPIC", "-std=gnu11", f"-I{NODE_INCLUDE}",
           "-DNODE_GYP_MODULE_NAME=flowhip", os.path.join(NODE_DIR, "flowhip_napi.c"),
           "-o", tmp, f"-L{HERE}", "-lflowhip", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"N-API addon build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, ADDON)
    return ADDON


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
