"""Engine bound to the CPU emulation of the kernels (tests/cpu_sim/_build/<flavor>/libsdim_sim.so).

Same C ABI and Python face as sdim_amd.engine.Engine; every kernel launch runs on
the host (tests/cpu_sim/simt.cpp).  Test infrastructure only.
"""
import ctypes
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from sdim_amd import _lib  # noqa: E402
from sdim_amd.engine import Engine  # noqa: E402

from tests.cpu_sim import build_sim  # noqa: E402


class SimEngine(Engine):
    def __init__(self, flavor="plain", **options):
        self.L = _lib.load(ensure_built(flavor))
        self.ctx = ctypes.c_void_p()
        _lib.check(self.L.sdim_ctx_create(0, ctypes.byref(self.ctx)), self.L)
        # the gang kernel is not part of the CPU build
        self.set_option("gang_min_n", 0)
        for k, v in options.items():
            self.set_option(k, v)


def ensure_built(flavor):
    """Builds the flavor when its library is missing or older than a source it is built from."""
    lib = build_sim.lib_path(flavor)
    srcs = [s for _, s, _ in build_sim.units()]
    srcs += [os.path.join(build_sim.CSRC, h) for h in os.listdir(build_sim.CSRC) if h.endswith(".h")]
    srcs += [os.path.join(HERE, "hip", "hip_runtime.h")]
    if not os.path.exists(lib) or os.path.getmtime(lib) < max(os.path.getmtime(s) for s in srcs):
        build_sim.build(flavor)
    return lib


def run_subprocess(flavor, args, timeout=1800, extra_env=None):
    """Runs tests/cpu_sim/run_sim.py with `args` under the flavor's sanitizer runtime."""
    ensure_built(flavor)
    env = dict(os.environ)
    rt = build_sim.sanitizer_runtime(flavor)
    if rt:
        env["LD_PRELOAD"] = rt + (":" + env["LD_PRELOAD"] if env.get("LD_PRELOAD") else "")
    if flavor == "asan":
        env.setdefault("ASAN_OPTIONS", "detect_leaks=0:abort_on_error=1:halt_on_error=1")
        env.setdefault("UBSAN_OPTIONS", "print_stacktrace=1:halt_on_error=1")
    if flavor == "tsan":
        env.setdefault("TSAN_OPTIONS", "halt_on_error=1:report_signal_unsafe=0:second_deadlock_stack=1")
    if extra_env:
        env.update(extra_env)
    return subprocess.run([sys.executable, os.path.join(HERE, "run_sim.py"), "--flavor", flavor, *args],
                          capture_output=True, text=True, timeout=timeout, env=env, cwd=ROOT)
