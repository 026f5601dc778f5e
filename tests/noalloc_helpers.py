"""Checks of the allocation-free entry point (sdim_simulate_noalloc) shared by its GPU test
and the CPU build: its records must equal those of the existing entry point (sdim_submit,
itself checked against the oracle) on the same jobs, on two calls in a row into one
buffer, the second with a smaller input."""
import ctypes

import numpy as np

from sdim_amd import _lib

from parity_helpers import clifford_job, rand_job


def noalloc_records(eng, jobs, out_ptr, capacity, to_host):
    """Runs the jobs through sdim_simulate_noalloc into the device buffer `out_ptr`; returns
    the bytes needed and each job's records as (value, deterministic) arrays [n_meas, shots]."""
    arr = (_lib.SdimJob * len(jobs))()
    for i, j in enumerate(jobs):
        arr[i] = j.to_c()
    need = ctypes.c_uint64(0)
    _lib.check(eng.L.sdim_simulate_noalloc(eng.ctx, arr, len(jobs), None, 0, ctypes.byref(need)), eng.L)
    if out_ptr is None:
        return need.value, None
    _lib.check(eng.L.sdim_simulate_noalloc(eng.ctx, arr, len(jobs), out_ptr, capacity, ctypes.byref(need)), eng.L)
    host = to_host(out_ptr, need.value)
    out, off = [], 0
    for j in jobs:
        m = int(np.isin(np.asarray(j.ir)[:, 0], (14, 15, 16)).sum())
        b = host[off:off + m * j.shots].reshape(m, j.shots)
        out.append(((b & 0x7F).astype(np.int64), (b >> 7).astype(bool)))
        off = (off + m * j.shots + 255) & ~255
    return need.value, out


def reference_records(eng, jobs):
    tk = eng.run(jobs)
    out = [(tk.records(i)['measurement_value'], tk.records(i)['deterministic']) for i in range(len(jobs))]
    tk.release()
    return out


def check_noalloc_against_submit(eng, rng, alloc, to_host, fill):
    """alloc(nbytes) -> device pointer; to_host(ptr, nbytes) -> uint8 array; fill(ptr, value, nbytes)."""
    big = [rand_job(rng, int(rng.integers(4, 14)), d, int(rng.integers(20, 80)), shots=int(rng.integers(1, 9)),
                    stream=i) for i, d in enumerate((2, 3, 5, 7, 11))]
    big += [clifford_job(rng, n, d, 3 * n, shots=4, stream=10 + i, noise_p=0.05)
            for i, (n, d) in enumerate(((40, 2), (90, 3), (230, 5), (300, 7)))]
    small = [big[1], big[6]]
    need_big, _ = noalloc_records(eng, big, None, 0, to_host)
    need_small, _ = noalloc_records(eng, small, None, 0, to_host)
    assert 0 < need_small < need_big
    ptr = alloc(need_big)
    for jobs, need in ((big, need_big), (small, need_small)):
        fill(ptr, 0xFF, need_big)           # stale bytes must not survive
        got_need, got = noalloc_records(eng, jobs, ptr, need_big, to_host)
        assert got_need == need
        ref = reference_records(eng, jobs)
        for k, ((v, d), (rv, rd)) in enumerate(zip(got, ref)):
            assert np.array_equal(v, rv) and np.array_equal(d, rd), f"job {k} (n={jobs[k].n}, d={jobs[k].d}) differs"
    # a buffer one byte short and oversized metadata are refused with a clear error
    arr = (_lib.SdimJob * len(big))()
    for i, j in enumerate(big):
        arr[i] = j.to_c()
    rc = eng.L.sdim_simulate_noalloc(eng.ctx, arr, len(big), ptr, need_big - 1, None)
    assert rc == -1 and b"output buffer holds" in eng.L.sdim_last_error()
    huge = [clifford_job(rng, 30, 3, 600_000, shots=1, stream=99)]    # > 8 MiB of lowered ops
    arr1 = (_lib.SdimJob * 1)()
    arr1[0] = huge[0].to_c()
    need = ctypes.c_uint64(0)
    rc = eng.L.sdim_simulate_noalloc(eng.ctx, arr1, 1, ptr, need_big, ctypes.byref(need))
    assert rc == -3 and b"allocation-free path has" in eng.L.sdim_last_error(), (rc, eng.L.sdim_last_error())
