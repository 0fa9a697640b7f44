"""Build the engine library ``mapol_amd/libbatchpf.so`` for gfx950.

One hipcc invocation compiles the kernels, the C-ABI, the host-side symbolic
analysis and the MATPOWER reader into a single shared library.  Packed FP32 instructions are
disabled for all device code (``-packed-fp32-ops``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libbatchpf.so")
ARCH = os.environ.get("BPF_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc",
                 shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def command(out: str = LIB):
    srcs = [os.path.join(CSRC, "engine.hip"), os.path.join(CSRC, "symbolic.cpp"),
            os.path.join(CSRC, "matpower.cpp")]
    return [hipcc(), "-O3", "-std=c++17", "-fPIC", "-shared",
            f"--offload-arch={ARCH}", "-Wno-unused-result", "-Wno-unused-value",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I", os.path.join(os.path.dirname(HERE), "include"),
            "-o", out] + srcs


def build(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
