"""Checks of the kernel's tableau snapshots (engine option "snapshots") shared by the GPU test
and the CPU build: snapshot e of mode 1 (at measurement event e, before RESET's fix-up) and
snapshot k of mode 2 (after instruction k) equal the final tableau of the circuit cut right
there (with the cut RESET turned into M for mode 1), every shot, batched and plain."""
import numpy as np

from parity_helpers import rand_job


def _prefix_job(job, k, reset_as_measure):
    from sdim_amd.engine import Job
    ir = np.asarray(job.ir)[:k + 1].copy()
    if reset_as_measure and ir[k, 0] == 16:
        ir[k, 0] = 14
    ch = None if job.noise_channel is None else np.asarray(job.noise_channel)[:k + 1]
    pr = None if job.noise_prob is None else np.asarray(job.noise_prob)[:k + 1]
    return Job(job.n, job.d, ir, shots=job.shots, seed=job.seed, stream=job.stream, noise_channel=ch, noise_prob=pr)


def check_snapshots(eng, rng, count=4):
    for d in (2, 3, 5, 7, 11):
        jobs = [rand_job(rng, int(rng.integers(2, 10)), d, int(rng.integers(8, 40)), shots=3, stream=i)
                for i in range(count)]
        for mode in ("batched", "plain"):
            for snap in (1, 2):
                eng.set_option("snapshots", snap)
                tk = eng.run(jobs, mode=mode)
                eng.set_option("snapshots", 0)
                for ji, job in enumerate(jobs):
                    ir = np.asarray(job.ir)
                    cuts = ([k for k in range(len(ir)) if ir[k, 0] in (14, 15, 16)] if snap == 1
                            else list(range(len(ir))))
                    for idx, k in enumerate(cuts):
                        ref = eng.run([_prefix_job(job, k, snap == 1)], mode=mode)
                        for s in range(job.shots):
                            for a, b in zip(tk.snapshot(ji, idx, s), ref.tableau(0, s)):
                                assert np.array_equal(a, b), (d, mode, snap, ji, idx, s)
                        ref.release()
                tk.release()
