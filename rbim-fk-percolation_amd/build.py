#!/usr/bin/env python3
"""Build librbim_hip.so in-tree for gfx950 (explicit hipcc, no JIT cache).

    python rbim-fk-percolation_amd/build.py

The library lands next to the Python package (``rbim/librbim_hip.so``), so it
travels with the source tree and is what ``rbim._hip`` loads.
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCES = ["csrc/lds_engine.hip", "csrc/tiled_engine.hip", "csrc/fused_engine.hip", "csrc/capi.hip"]
OUTPUT = os.path.join(HERE, "rbim", "librbim_hip.so")
ARCH = os.environ.get("RBIM_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Wall",
    # packed FP32 instructions stay off in all device code of this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # The kernels' atomics on a wavefront-uniform address are made by one lane
    # (a cluster's first-lane count) or by the few lanes with a wrap bit; the
    # compiler's wavefront-reduction rewrite of such atomics only adds
    # instructions to them.
    "-mllvm", "-amdgpu-atomic-optimizer-strategy=None",
]


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "hipcc")
    tmp = OUTPUT + ".tmp"
    cmd = [hipcc, *FLAGS, "-o", tmp, *[os.path.join(HERE, s) for s in SOURCES]]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"hipcc failed with status {proc.returncode}")
    os.replace(tmp, OUTPUT)
    return OUTPUT


if __name__ == "__main__":
    print(build(verbose=True))
