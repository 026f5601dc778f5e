"""Build the CPU emulation of the engine: the kernel sources of sdim_amd/csrc compiled
for the host against the stand-in hip/hip_runtime.h of this directory.

    python tests/cpu_sim/build_sim.py [plain|asan|tsan ...]

Flavors: plain (-O2), asan (AddressSanitizer + UndefinedBehaviorSanitizer: every LDS
and device-memory access checked against its allocation), tsan (ThreadSanitizer: the
threads of a workgroup must communicate through barriers, atomics or wave exchanges).
Output: tests/cpu_sim/_build/<flavor>/libsdim_sim.so.  Test infrastructure only: the
product library is sdim_amd/libsdim_hip.so and never loads this one.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "sdim_amd", "csrc")
CLANG = os.environ.get("SDIM_SIM_CXX", "/opt/rocm/llvm/bin/clang++")
FLAGS = {
    "plain": ["-O2"],
    "asan": ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
             "-fno-sanitize-recover=undefined"],
    "tsan": ["-O1", "-g", "-fsanitize=thread"],
}


def units():
    u = [(f"batch_{nt}", os.path.join(CSRC, "batch.hip"), [f"-DSDIM_BATCH_NT={nt}"]) for nt in (64, 128, 256, 512, 1024)]
    u += [("plain", os.path.join(CSRC, "plain.hip"), []), ("frame", os.path.join(CSRC, "frame.hip"), []),
          ("noalloc", os.path.join(CSRC, "noalloc.hip"), []),
          ("runtime", os.path.join(CSRC, "runtime.cpp"), []), ("simt", os.path.join(HERE, "simt.cpp"), []),
          ("gang_stub", os.path.join(HERE, "gang_stub.cpp"), []),
          ("selftest", os.path.join(HERE, "selftest.cpp"), [])]
    return u


def lib_path(flavor):
    return os.path.join(HERE, "_build", flavor, "libsdim_sim.so")


def sanitizer_runtime(flavor):
    """Path of the sanitizer runtime to LD_PRELOAD into an uninstrumented host (python)."""
    name = {"asan": "libclang_rt.asan-x86_64.so", "tsan": "libclang_rt.tsan-x86_64.so"}.get(flavor)
    if not name:
        return None
    p = subprocess.run([CLANG, f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    if not os.path.exists(p):
        p = subprocess.run([CLANG, "-print-resource-dir"], capture_output=True, text=True).stdout.strip()
        p = os.path.join(p, "lib", "x86_64-unknown-linux-gnu", name)
    return p


def build(flavor="plain", verbose=False):
    out_dir = os.path.join(HERE, "_build", flavor)
    os.makedirs(out_dir, exist_ok=True)
    common = ["-x", "c++", "-std=c++17", "-fPIC", "-pthread", "-I", HERE, "-I", os.path.join(ROOT, "include"),
              "-I", CSRC, "-DSDIM_WAVES_PER_EU(n)=", "-Wno-unknown-attributes", "-Wno-ignored-attributes", "-Wno-unused-command-line-argument"]
    cmds = [[CLANG, *common, *FLAGS[flavor], *defs, "-c", src, "-o", os.path.join(out_dir, name + ".o")]
            for name, src, defs in units()]

    def run(c):
        if verbose:
            print(" ".join(c), flush=True)
        return subprocess.run(c, capture_output=True, text=True)

    with ThreadPoolExecutor(max_workers=os.cpu_count() or 4) as ex:
        res = list(ex.map(run, cmds))
    bad = [r for r in res if r.returncode]
    for r in bad:
        sys.stderr.write(r.stdout + r.stderr)
    if bad:
        raise RuntimeError(f"CPU emulation build ({flavor}) failed")
    tmp = lib_path(flavor) + ".tmp"
    link = [CLANG, "-shared", "-pthread", *FLAGS[flavor], *[os.path.join(out_dir, n + ".o") for n, _, _ in units()],
            "-o", tmp]
    if flavor != "plain":
        link.append("-shared-libsan")
    r = run(link)
    if r.returncode:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"CPU emulation link ({flavor}) failed")
    os.replace(tmp, lib_path(flavor))   # a process that has the old library mapped keeps it
    return lib_path(flavor)


if __name__ == "__main__":
    for f in (sys.argv[1:] or ["plain"]):
        print(build(f, verbose=True))
