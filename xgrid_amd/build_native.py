"""Build libxgrid_hip.so (host runtime + probe kernels) for gfx950 in-tree.

Run as ``python -m xgrid_amd.build_native``.  The generated stencil kernels are
compiled later, at first call, with hipRTC (``xgrid_amd/rtc.py``).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libxgrid_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

# Packed FP32 instructions stay off in every code object this project builds.
NO_PACKED_FP32 = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc() -> str:
    return os.path.join(ROCM, "bin", "hipcc")


def build_runtime(verbose: bool = False) -> str:
    src = os.path.join(CSRC, "xgrid_runtime.cpp")
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-x", "hip", "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-ffp-contract=off", *NO_PACKED_FP32, f"-I{INCLUDE}", f"-I{CSRC}", src,
           "-o", tmp, f"-L{ROCM}/lib", "-lhiprtc", "-lamdhip64", f"-Wl,-rpath,{ROCM}/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_runtime(verbose="-v" in sys.argv))
