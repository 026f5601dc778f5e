"""The engine's own kernels compiled for the CPU (tests/cpu_sim): parity with the oracle,
every LDS and device-memory access checked by AddressSanitizer, and the threads of each
workgroup checked by ThreadSanitizer for communication outside barriers, atomics and wave
exchanges.  Each case runs in a subprocess (the sanitizer runtimes must be preloaded).

The checkers themselves are pinned first: a kernel with a planted LDS race must be
reported by ThreadSanitizer, one with a write one byte past its dynamic LDS by
AddressSanitizer, and their race-free / in-bounds twins must not be.
"""
import os
import shutil

import pytest

from tests.cpu_sim import build_sim, sim_engine

pytestmark = pytest.mark.skipif(not os.path.exists(build_sim.CLANG) or not shutil.which("make"),
                                reason="needs the ROCm clang++ for the CPU build of the kernels")

TSAN_REPORT = "WARNING: ThreadSanitizer"
ASAN_REPORT = "ERROR: AddressSanitizer"


def _run(flavor, *args, timeout=1200, env=None):
    r = sim_engine.run_subprocess(flavor, list(args), timeout=timeout, extra_env=env)
    return r, r.stdout + r.stderr


@pytest.mark.parametrize("which,flavor,expect", [
    (0, "tsan", None), (1, "tsan", TSAN_REPORT), (3, "tsan", None),
    (0, "asan", None), (2, "asan", ASAN_REPORT), (3, "asan", None),
])
def test_checkers_catch_planted_bugs(which, flavor, expect):
    r, out = _run(flavor, "--case", "selftest", "--which", str(which), timeout=300)
    if expect is None:
        assert r.returncode == 0 and TSAN_REPORT not in out and ASAN_REPORT not in out, out[-3000:]
    else:
        assert r.returncode != 0 and expect in out, out[-3000:]


def test_cpu_build_matches_oracle():
    """Random circuits (every gate id, noise, M / M_X / RESET, d in {2,3,5,7,11}) through the
    batched path (LDS- and HBM-resident items) and the plain path, records and tableaux."""
    r, out = _run("plain", "--case", "random", "--seed", "11", "--count", "10", "--hbm", timeout=600)
    assert r.returncode == 0, out[-4000:]


def test_cpu_build_snapshots_and_noalloc():
    """Tableau snapshots (per measurement event, per instruction) against prefix runs, and the
    allocation-free entry point against sdim_submit, on the CPU build."""
    for case in ("snapshots", "noalloc"):
        r, out = _run("plain", "--case", case, "--count", "2", timeout=600)
        assert r.returncode == 0, out[-4000:]


def test_serial_layers_match_oracle():
    """Gate segments with runs of small layers applied serially (serial_max): random circuits
    in LDS and HBM, and uneven-batch circuits with every run merged, against the oracle;
    the random circuits again under AddressSanitizer."""
    r, out = _run("plain", "--case", "random", "--seed", "21", "--count", "8", "--hbm", "--opt", "serial_max=8")
    assert r.returncode == 0, out[-4000:]
    r, out = _run("plain", "--case", "uneven", "--first", "3", "--stride", "512", "--check", "8",
                  "--opt", "serial_max=1000")
    assert r.returncode == 0, out[-4000:]
    r, out = _run("asan", "--case", "random", "--seed", "24", "--count", "3", "--hbm", "--opt", "serial_max=6")
    assert r.returncode == 0 and ASAN_REPORT not in out, out[-6000:]


def test_two_qudit_gates_on_one_qudit_match_oracle():
    """SUM, SUM^-1, CZ, CZ^-1 and SWAP whose two qudits coincide (the reference accepts them; CZ
    and SWAP then follow the reference's literal words), with serial layers (rows in registers
    for tableaux of at most 16 qudits) and without, in LDS and HBM, batched and plain."""
    r, out = _run("plain", "--case", "random", "--seed", "31", "--count", "6", "--hbm", "--same", "0.4",
                  "--opt", "serial_max=8")
    assert r.returncode == 0, out[-4000:]
    r, out = _run("plain", "--case", "random", "--seed", "32", "--count", "6", "--same", "0.4",
                  "--opt", "serial_max=0")
    assert r.returncode == 0, out[-4000:]
    # one-wave HBM-resident workgroups: the register form of serial layers on LDS column tiles
    r, out = _run("plain", "--case", "random", "--seed", "41", "--count", "6", "--same", "0.2",
                  "--opt", "hbm_threads=64", "--opt", "allow_lds=0", "--opt", "serial_max=8")
    assert r.returncode == 0, out[-4000:]


def test_address_sanitizer_random_circuits():
    r, out = _run("asan", "--case", "random", "--seed", "12", "--count", "4", "--hbm")
    assert r.returncode == 0 and ASAN_REPORT not in out, out[-6000:]


def test_address_sanitizer_uneven_batch_sample():
    """Every 32nd circuit of the uneven batch with the device's launch plan (all launch
    classes: LDS-resident 128..1024 threads, the HBM-resident class with column tiles and
    LDS phases), under AddressSanitizer; two circuits also against the oracle."""
    r, out = _run("asan", "--case", "uneven", "--first", "5", "--stride", "32", "--check", "2")
    assert r.returncode == 0 and ASAN_REPORT not in out, out[-6000:]


def test_thread_sanitizer_random_circuits():
    r, out = _run("tsan", "--case", "random", "--seed", "13", "--count", "4", "--hbm",
                  env={"SDIM_SIM_THREADS": "4"})
    assert r.returncode == 0 and TSAN_REPORT not in out, out[-6000:]


def test_thread_sanitizer_uneven_batch_sample():
    r, out = _run("tsan", "--case", "uneven", "--first", "7", "--stride", "256", "--check", "0",
                  env={"SDIM_SIM_THREADS": "4"})
    assert r.returncode == 0 and TSAN_REPORT not in out, out[-6000:]


def test_list_overflow_is_reported_not_written():
    """With the LDS lists capped at one entry every measurement overflows them: the writes
    are dropped (AddressSanitizer sees no access out of bounds) and the ticket fails."""
    r, out = _run("asan", "--case", "uneven", "--first", "3", "--stride", "512", "--check", "0",
                  "--opt", "list_cap=1", "--expect-error", "capacity check failed", timeout=600)
    assert r.returncode == 0 and ASAN_REPORT not in out, out[-6000:]
