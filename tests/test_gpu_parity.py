"""GPU parity: the HIP engine against the oracle, bit-exact, through the C ABI.

Compares measurement records (value, deterministic flag, qudit, round, shot) and
the final tableau of every shot with the C restatement of the oracle
(oracle/sdim_oracle.c, itself pinned by tests/test_oracle_pinning.py) on the
same Philox key.
"""
import numpy as np
import pytest

from oracle import c_oracle, philox

pytestmark = pytest.mark.gpu


def rand_job(rng, n, d, depth, meas=0.15, noise=0.1, shots=3, stream=0):
    from sdim_amd.engine import Job
    ops, ch, pr = [], [], []
    for _ in range(depth):
        r = rng.random()
        if r < meas:
            ops.append((int(rng.choice([14, 15, 16])), int(rng.integers(n)), -1)); ch.append(0); pr.append(0.0)
        elif r < meas + noise:
            ops.append((17, int(rng.integers(n)), -1)); ch.append(int(rng.integers(3))); pr.append(float(rng.random()))
        elif r < 0.65 or n < 2:
            ops.append((int(rng.integers(0, 9)), int(rng.integers(n)), -1)); ch.append(0); pr.append(0.0)
        else:
            a, b = rng.choice(n, 2, replace=False)
            ops.append((int(rng.integers(9, 14)), int(a), int(b))); ch.append(0); pr.append(0.0)
    ir = np.asarray(ops, dtype=np.int64).reshape(-1, 3)
    return Job(n, d, ir, shots=shots, seed=int(rng.integers(1 << 62)), stream=stream,
               noise_channel=np.asarray(ch, np.int32), noise_prob=np.asarray(pr, np.float64))


def oracle_of(job, keep_shots):
    noise = {}
    ir = np.asarray(job.ir).reshape(-1, 3)
    for k in range(ir.shape[0]):
        if ir[k, 0] == 17:
            noise[k] = (int(job.noise_channel[k]), philox.keep_threshold53(float(job.noise_prob[k])))
    return c_oracle.simulate(job.n, job.d, ir, shots=job.shots, seed=job.seed, stream=job.stream,
                             noise=noise, keep_shots=keep_shots)


def check_ticket(tk, jobs, tableaux=True):
    for ji, job in enumerate(jobs):
        rec_o, tabs_o = oracle_of(job, job.shots if tableaux else 0)
        rec = tk.records(ji)                      # [M, shots]
        got = np.stack([rec['qudit_index'], rec['meas_round'], rec['shot'],
                        rec['deterministic'].astype(np.int64), rec['measurement_value']], axis=-1)
        got = np.transpose(got, (1, 0, 2))        # [shots, M, 5]
        assert np.array_equal(got, rec_o), f"records differ for job {ji} (n={job.n}, d={job.d})"
        if tableaux:
            for s in range(job.shots):
                t = tk.tableau(ji, s)
                for name, a, b in zip(("x", "z", "ph", "dx", "dz", "dph"), t, tabs_o[s]):
                    assert np.array_equal(a, b), f"tableau {name} differs, job {ji} shot {s} (n={job.n}, d={job.d})"


@pytest.fixture(scope="module")
def engine():
    from sdim_amd.engine import Engine
    return Engine()


@pytest.mark.parametrize("mode", ["batched", "plain"])
@pytest.mark.parametrize("d", [2, 3, 5, 7, 11])
def test_random_circuits_small(engine, mode, d):
    rng = np.random.default_rng(100 * d + (mode == "plain"))
    jobs = [rand_job(rng, int(rng.integers(1, 12)), d, int(rng.integers(1, 60)), stream=i) for i in range(12)]
    print(f"[parity] mode={mode} d={d} jobs={len(jobs)}", flush=True)
    tk = engine.run(jobs, mode=mode)
    check_ticket(tk, jobs)
    tk.release()
