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


SOURCES = ["xgrid_runtime.cpp", "xgrid_reduce.cpp", "xgrid_batch.cpp"]


def build_runtime(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-x", "hip", "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-ffp-contract=off", *NO_PACKED_FP32, f"-I{INCLUDE}", f"-I{CSRC}", *srcs,
           "-o", tmp, f"-L{ROCM}/lib", "-lhiprtc", "-lamdhip64", f"-Wl,-rpath,{ROCM}/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    # every symbol of our own must resolve inside the library (a shared object links with holes)
    nm = subprocess.run(["nm", "-D", "-C", "--undefined-only", tmp], capture_output=True, text=True, check=True)
    own = [ln for ln in nm.stdout.splitlines() if "xgi::" in ln or " xg_" in ln]
    if own:
        raise RuntimeError("libxgrid_hip.so has unresolved symbols:\n" + "\n".join(own))
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_runtime(verbose="-v" in sys.argv))
