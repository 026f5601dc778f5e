"""Build the HIP engine library (gfx950) in-tree: sdim_amd/libsdim_hip.so.

    python -m sdim_amd.build            # compile
    python -m sdim_amd.build --print    # show the command only
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsdim_hip.so")
SOURCES = ["engine.hip", "frame.hip", "runtime.cpp"]
HEADERS = ["engine.h", "sdim_types.h", "sdim_arith.h"]


def command(out=LIB):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    return [hipcc, "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
            # packed FP32 VALU ops are never wanted in this integer engine
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument",
            "-I", os.path.join(HERE, "..", "include"),
            *[os.path.join(CSRC, s) for s in SOURCES], "-o", out]


def build(verbose=False):
    tmp = LIB + ".tmp"
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libsdim_hip.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    if "--print" in sys.argv:
        print(" ".join(command()))
    else:
        print(build(verbose=True))
