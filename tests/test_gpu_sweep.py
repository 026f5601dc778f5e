"""GPU parity sweep: random circuits of every gate id and dimension at sizes around every launch
class limit (n from 1 to 300: small-tableau, LDS-resident and HBM-resident classes, gangs), with
noise, measurements, RESET and two-qudit gates on one qudit, under eleven engine configurations
and the plain path; records and final tableaux of every shot against the oracle.
"""
import numpy as np
import pytest

from parity_helpers import check_ticket, clifford_job, rand_job

pytestmark = pytest.mark.gpu

CONFIGS = [dict(), dict(allow_lds=0), dict(serial_max=16), dict(serial_max=0, allow_lds=0),
           dict(lds_bytes=32768, seg_batches=1), dict(interleave=2, allow_lds=0), dict(shots_per_item=3),
           dict(hbm_threads=512, hbm_per_cu=2, sweep_rows=4), dict(gang_min_n=100, gang_chunk=16, gang_max=8),
           dict(gang_min_n=100, gang_chunk=16, gang_max=4, gang_threads=256, serial_max=16), dict(mode="plain")]


@pytest.mark.parametrize("k", range(len(CONFIGS)))
def test_random_sizes_and_options(k):
    from sdim_amd.engine import Engine
    opts = dict(CONFIGS[k])
    mode = opts.pop("mode", "batched")
    rng = np.random.default_rng(9000 + k)
    jobs = []
    for i in range(60):
        d = int(rng.choice([2, 3, 5, 7, 11]))
        n = int(rng.choice([1, 2, 7, 15, 16, 17, 33, 64, 97, 150, 220, 300]))
        if n < 2 or rng.random() < 0.5:
            jobs.append(rand_job(rng, n, d, int(rng.integers(1, 4 * n + 10)), shots=int(rng.integers(1, 70)),
                                 stream=i, same=0.05))
        else:
            jobs.append(clifford_job(rng, n, d, 3 * n, shots=int(rng.integers(1, 20)), stream=i,
                                     noise_p=0.03 if i % 3 == 0 else 0.0))
    eng = Engine(**opts)
    print(f"[sweep] mode={mode} opts={opts} jobs={len(jobs)}", flush=True)
    tk = eng.run(jobs, mode=mode)
    check_ticket(tk, jobs)
    tk.release()
    eng.close()
