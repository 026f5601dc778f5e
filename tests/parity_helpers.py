"""Shared helpers of the GPU parity tests: random circuit jobs and the bit-exact
comparison of a ticket's records and final tableaux with the C oracle
(oracle/sdim_oracle.c, pinned by tests/test_oracle_pinning.py) on the same
Philox key."""
import numpy as np

from oracle import c_oracle, philox


def rand_job(rng, n, d, depth, meas=0.15, noise=0.1, shots=3, stream=0, same=0.0):
    """Random circuit over every gate id; `same` is the share of two-qudit gates whose two qudits
    coincide (the reference accepts them)."""
    from sdim_amd.engine import Job
    ops, ch, pr = [], [], []
    for _ in range(depth):
        r = rng.random()
        if r < meas:
            ops.append((int(rng.choice([14, 15, 16])), int(rng.integers(n)), -1)); ch.append(0); pr.append(0.0)
        elif r < meas + noise:
            ops.append((17, int(rng.integers(n)), -1)); ch.append(int(rng.integers(3))); pr.append(float(rng.random()))
        elif same > 0 and r < meas + noise + same * (1 - meas - noise):
            a = int(rng.integers(n))
            ops.append((int(rng.integers(9, 14)), a, a)); ch.append(0); pr.append(0.0)
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


def check_job(tk, ji, job, tableaux=True, shots=None):
    """Records of every shot of job `ji` of ticket `tk`, and the final tableau of the
    shots in `shots` (all by default), bit-exact against the C oracle."""
    keep = (job.shots if shots is None else max(shots) + 1) if tableaux else 0
    rec_o, tabs_o = oracle_of(job, keep)
    rec = tk.records(ji)                      # [M, shots]
    got = np.stack([rec['qudit_index'], rec['meas_round'], rec['shot'],
                    rec['deterministic'].astype(np.int64), rec['measurement_value']], axis=-1)
    got = np.transpose(got, (1, 0, 2))        # [shots, M, 5]
    assert np.array_equal(got, rec_o), f"records differ for job {ji} (n={job.n}, d={job.d})"
    if tableaux:
        for s in (range(job.shots) if shots is None else shots):
            t = tk.tableau(ji, s)
            for name, a, b in zip(("x", "z", "ph", "dx", "dz", "dph"), t, tabs_o[s]):
                assert np.array_equal(a, b), f"tableau {name} differs, job {ji} shot {s} (n={job.n}, d={job.d})"


def check_ticket(tk, jobs, tableaux=True):
    for ji, job in enumerate(jobs):
        check_job(tk, ji, job, tableaux)


def clifford_job(rng, n, d, depth, shots, stream, meas_rounds=1, noise_p=0.0):
    """Random {H,P,SUM,X,Z,H_INV,P_INV,SUM_INV,CZ,SWAP} circuit + measure-all (+ optional N1 after gates)."""
    from sdim_amd.engine import Job
    gates = np.array([5, 7, 9, 1, 3, 6, 8, 10, 11, 13])
    ops, ch, pr = [], [], []
    for _ in range(depth):
        g = int(rng.choice(gates))
        if g in (9, 10, 11, 13):
            a, b = rng.choice(n, 2, replace=False)
            ops.append((g, int(a), int(b)))
        else:
            ops.append((g, int(rng.integers(n)), -1))
        ch.append(0); pr.append(0.0)
        if noise_p:
            ops.append((17, ops[-1][1], -1)); ch.append(int(rng.integers(3))); pr.append(noise_p)
    for _ in range(meas_rounds):
        for q in range(n):
            ops.append((14 if rng.random() < 0.7 else 16, q, -1)); ch.append(0); pr.append(0.0)
    return Job(n, d, np.asarray(ops, np.int64), shots=shots, seed=int(rng.integers(1 << 62)), stream=stream,
               noise_channel=np.asarray(ch, np.int32), noise_prob=np.asarray(pr, np.float64))


def records_array(rec):
    """MEASUREMENT_DTYPE records (events, shots) -> int64 [shots, events, 5] (the oracle's layout)."""
    got = np.stack([rec['qudit_index'], rec['meas_round'], rec['shot'],
                    rec['deterministic'].astype(np.int64), rec['measurement_value']], axis=-1)
    return np.ascontiguousarray(np.transpose(got, (1, 0, 2)), dtype=np.int64)


def digest_records(arr):
    """16-byte SHA-256 prefix of int64 records [shots, events, 5]."""
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(arr, dtype=np.int64).tobytes()).digest()[:16]


def digest_tableau(tab):
    """16-byte SHA-256 prefix of a tableau (x, z, ph, dx, dz, dph) as int64 in the reference layout."""
    import hashlib
    h = hashlib.sha256()
    for a in tab:
        h.update(np.ascontiguousarray(a, dtype=np.int64).tobytes())
    return h.digest()[:16]
