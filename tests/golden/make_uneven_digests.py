"""Writes tests/golden/uneven_batch_digests.npz: oracle digests of the whole uneven batch.

    python tests/golden/make_uneven_digests.py [--procs N]

For each of the 4096 circuits of sdim_amd.workloads.uneven_batch() (n ~ U[32,512],
d ~ U{2,3,5,7}, 10n gates from {H,P,SUM,X,Z} + measure-all, 8 shots) the C oracle
(oracle/sdim_oracle.c, pinned against the literal restatement of the reference by
tests/test_oracle_pinning.py) is run on the circuit's own Philox key, and two
16-byte SHA-256 prefixes are stored: of all 8 shots' records and of shot 0's final
tableau (digests as defined in tests/parity_helpers.py).  Runs on the CPU in this
container (a few minutes on 8 cores); the GPU test compares every circuit against
these digests.
"""
import argparse
import os
import sys
import time
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
OUT = os.path.join(HERE, "uneven_batch_digests.npz")


def one(i):
    from oracle import c_oracle
    from parity_helpers import digest_records, digest_tableau
    from sdim_amd import workloads as W
    j = W.uneven_batch().jobs[i] if not hasattr(one, "wl") else one.wl[i]
    rec, tabs = c_oracle.simulate(j.n, j.d, np.asarray(j.ir), shots=j.shots, seed=j.seed, stream=j.stream,
                                  keep_shots=1)
    return i, digest_records(rec), digest_tableau(tabs[0])


def init():
    from sdim_amd import workloads as W
    one.wl = W.uneven_batch().jobs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--procs", type=int, default=os.cpu_count() or 4)
    a = ap.parse_args()
    from sdim_amd import workloads as W
    jobs = W.uneven_batch().jobs
    N = len(jobs)
    rec = np.zeros((N, 16), np.uint8)
    tab = np.zeros((N, 16), np.uint8)
    t0 = time.time()
    # largest circuits first so the pool finishes together
    order = sorted(range(N), key=lambda i: -jobs[i].n)
    with Pool(a.procs, initializer=init) as pool:
        for k, (i, r, t) in enumerate(pool.imap_unordered(one, order, chunksize=4)):
            rec[i] = np.frombuffer(r, np.uint8)
            tab[i] = np.frombuffer(t, np.uint8)
            if k % 256 == 0:
                print(f"{k}/{N} circuits, {time.time() - t0:.0f} s", flush=True)
    np.savez(OUT, n=np.array([j.n for j in jobs], np.int32), d=np.array([j.d for j in jobs], np.int32),
             n_ops=np.array([np.asarray(j.ir).shape[0] for j in jobs], np.int32),
             seed=np.array([j.seed for j in jobs], np.uint64), stream=np.array([j.stream for j in jobs], np.uint64),
             records=rec, tableau0=tab)
    print(f"wrote {OUT} ({N} circuits, {time.time() - t0:.0f} s)")


if __name__ == "__main__":
    main()
