"""Build the HIP engine library (gfx950) in-tree: sdim_amd/libsdim_hip.so.

    python -m sdim_amd.build            # compile every object (in parallel) and link
    python -m sdim_amd.build --print    # show the commands only

The persistent batch kernel is compiled once per workgroup size and the gang
kernel once per dimension (one object each, built concurrently); the objects
go to sdim_amd/_build/ and are linked into one shared library.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsdim_hip.so")
HEADERS = ["engine.h", "sdim_types.h", "sdim_arith.h", "tableau_dev.h"]
BATCH_SIZES = (64, 128, 256, 512, 1024)
GANG_DIMS = (2, 3, 5, 7, 11)


def units():
    """(object name, source, extra defines) of every translation unit."""
    # the one-wave kernel (launch class of small tableaux) gets 256 VGPRs, two waves per SIMD: its
    # serial gate layers keep 2 x 16 rows of a column word in registers
    out = [(f"batch_{nt}", "batch.hip", [f"-DSDIM_BATCH_NT={nt}"] + (["-DSDIM_BATCH_WAVES=2"] if nt == 64 else []))
           for nt in BATCH_SIZES]
    out += [(f"gang_{d}", "gang.hip", [f"-DSDIM_GANG_D={d}"]) for d in GANG_DIMS]
    out += [("plain", "plain.hip", []), ("frame", "frame.hip", []), ("noalloc", "noalloc.hip", []),
            ("runtime", "runtime.cpp", [])]
    return out


def _hipcc():
    return os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def compile_command(name, src, defs):
    return [_hipcc(), "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-c",
            # packed FP32 VALU ops are never wanted in this integer engine
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument",
            "-I", os.path.join(HERE, "..", "include"), *defs,
            os.path.join(CSRC, src), "-o", os.path.join(OBJ, name + ".o")]


def link_command(out=LIB):
    return [_hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC",
            *[os.path.join(OBJ, name + ".o") for name, _, _ in units()], "-o", out]


def build(verbose=False, jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    cmds = [compile_command(*u) for u in units()]
    if verbose:
        for c in cmds:
            print(" ".join(c), flush=True)

    def run(cmd):
        return subprocess.run(cmd, capture_output=True, text=True)

    # the slowest objects first, so the pool finishes together
    order = sorted(range(len(cmds)), key=lambda i: 0 if units()[i][1] in ("batch.hip", "gang.hip") else 1)
    with ThreadPoolExecutor(max_workers=jobs or min(16, os.cpu_count() or 4)) as ex:
        results = dict(zip(order, ex.map(run, [cmds[i] for i in order])))
    failed = [i for i in range(len(cmds)) if results[i].returncode != 0]
    for i in failed:
        sys.stderr.write(results[i].stdout + results[i].stderr)
    if failed:
        raise RuntimeError("hipcc failed building " + ", ".join(units()[i][0] for i in failed))
    tmp = LIB + ".tmp"
    cmd = link_command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    r = run(cmd)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed linking libsdim_hip.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    if "--print" in sys.argv:
        for u in units():
            print(" ".join(compile_command(*u)))
        print(" ".join(link_command()))
    else:
        print(build(verbose=True))
