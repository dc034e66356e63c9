#!/usr/bin/env python3
"""Build librbim_hip.so in-tree for gfx950 (explicit hipcc, no JIT cache).

    python rbim-fk-percolation_amd/build.py

The library lands next to the Python package (``rbim/librbim_hip.so``), so it
travels with the source tree and is what ``rbim._hip`` loads.

Each source is compiled in two passes: the device pass (gfx950 code object,
with the device-only code-generation flags below) and the host pass, which
embeds that code object; the objects are then linked into one shared library.
Splitting the passes keeps the device target features off the host compiler,
which does not know them.
"""

from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCES = ["csrc/lds_engine.hip", "csrc/tiled_engine.hip", "csrc/fused_engine.hip", "csrc/capi.hip"]
OUTPUT = os.path.join(HERE, "rbim", "librbim_hip.so")
OBJDIR = os.path.join(HERE, "_build")
ARCH = os.environ.get("RBIM_OFFLOAD_ARCH", "gfx950")

COMMON = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall"]
DEVICE_ONLY = [
    # packed FP32 instructions stay off in all device code of this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # The kernels' atomics on a wavefront-uniform address are made by one lane
    # (a cluster's first-lane count) or by the few lanes with a wrap bit; the
    # compiler's wavefront-reduction rewrite of such atomics only adds
    # instructions to them.
    "-mllvm", "-amdgpu-atomic-optimizer-strategy=None",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"hipcc failed with status {proc.returncode}: {' '.join(cmd)}")


def _compile(hipcc, src, verbose):
    stem = os.path.splitext(os.path.basename(src))[0]
    dev = os.path.join(OBJDIR, stem + ".gpubin")
    obj = os.path.join(OBJDIR, stem + ".o")
    path = os.path.join(HERE, src)
    _run([hipcc, *COMMON, *DEVICE_ONLY, "--offload-device-only", "-c", "-o", dev, path], verbose)
    _run([hipcc, *COMMON, "--offload-host-only", "-Xclang", "-fcuda-include-gpubinary",
          "-Xclang", dev, "-c", "-o", obj, path], verbose)
    return obj


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "hipcc")
    os.makedirs(OBJDIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as pool:
        objs = list(pool.map(lambda s: _compile(hipcc, s, verbose), SOURCES))
    tmp = OUTPUT + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs], verbose)
    os.replace(tmp, OUTPUT)
    return OUTPUT


if __name__ == "__main__":
    print(build(verbose=True))
