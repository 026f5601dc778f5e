"""Synthetic workloads of the benchmark configurations, as IR arrays.

Gate choice follows ``generate_random_clifford_circuit`` (sdim/random_circuit.py:9-68):
uniform over a gate set, two-qudit gates on a uniformly random ordered pair
of distinct qudits, single-qudit gates on a uniform qudit, then measure-all
rounds.  A numpy Generator (PCG64) seeded per workload makes them reproducible.
"""
from dataclasses import dataclass

import numpy as np

from .engine import Job

H, H_INV, P, P_INV, CNOT, X, Z, M, RESET, N1 = 5, 6, 7, 8, 9, 1, 3, 14, 16, 17
TWO_QUDIT = {9, 10, 11, 12, 13}


def random_clifford_ir(rng, n, d, depth, gate_ids, measure_rounds=1):
    gate_ids = np.asarray(gate_ids, dtype=np.int64)
    g = gate_ids[rng.integers(len(gate_ids), size=depth)]
    q = rng.integers(n, size=depth)
    t = rng.integers(max(n - 1, 1), size=depth)
    t = t + (t >= q)
    two = np.isin(g, list(TWO_QUDIT))
    t = np.where(two, t, -1)
    body = np.stack([g, q, t], axis=1)
    meas = np.stack([np.full(n * measure_rounds, M), np.tile(np.arange(n), measure_rounds),
                     np.full(n * measure_rounds, -1)], axis=1)
    return np.concatenate([body, meas]).astype(np.int64)


def count_gate_shots(jobs):
    """Non-identity ops (including M/RESET/N1) times shots, summed over jobs."""
    tot = 0
    for j in jobs:
        ir = np.asarray(j.ir).reshape(-1, 3)
        tot += int((ir[:, 0] != 0).sum()) * int(j.shots)
    return tot


@dataclass
class Workload:
    name: str
    jobs: list
    description: str

    @property
    def gate_shots(self):
        return count_gate_shots(self.jobs)


def plain_small(seed=1, shots=1024):
    """d=2, n=24, 500 uniform {H, P, CNOT} then M on all qubits, 1024 shots."""
    rng = np.random.default_rng(seed)
    ir = random_clifford_ir(rng, 24, 2, 500, [H, P, CNOT])
    return Workload("plain_small_d2_n24", [Job(24, 2, ir, shots=shots, seed=seed, stream=0)],
                    "d=2 n=24 500 {H,P,CNOT} + M all, 1024 shots")


def repetition_code(seed=2, shots=4096, rounds=10, ring=17, p=1e-3):
    """d=3 ring of 17 data + 17 ancilla qutrits; per round SUM(d_i->a_i), SUM(d_{i+1}->a_i),
    ancilla M and RESET; after each SUM an X^a (channel 'f') and a Z^b (channel 'p') error
    on both touched qutrits with probability p."""
    n = 2 * ring
    ops, ch, pr = [], [], []

    def add(g, q, t=-1, c=0, prob=0.0):
        ops.append((g, q, t)); ch.append(c); pr.append(prob)

    for _ in range(rounds):
        for i in range(ring):
            a = ring + i
            for dq in (i, (i + 1) % ring):
                add(CNOT, dq, a)
                for q in (dq, a):
                    add(N1, q, -1, 1, p)
                    add(N1, q, -1, 2, p)
        for i in range(ring):
            add(M, ring + i)
        for i in range(ring):
            add(RESET, ring + i)
    ir = np.asarray(ops, dtype=np.int64)
    job = Job(n, 3, ir, shots=shots, seed=seed, stream=0,
              noise_channel=np.asarray(ch, np.int32), noise_prob=np.asarray(pr, np.float64))
    return Workload("repetition_d3_ring17", [job], "d=3 repetition ring 17+17, 10 rounds, p=1e-3, 4096 shots")


def uneven_batch(seed=3, circuits=4096, shots=8, n_lo=32, n_hi=512, dims=(2, 3, 5, 7), depth_per_qudit=10):
    """4096 random circuits, n ~ U[32,512], d ~ U{2,3,5,7}, 10n ops from {H,P,SUM,X,Z}, measure-all, 8 shots."""
    rng = np.random.default_rng(seed)
    jobs = []
    for c in range(circuits):
        n = int(rng.integers(n_lo, n_hi + 1))
        d = int(rng.choice(dims))
        ir = random_clifford_ir(rng, n, d, depth_per_qudit * n, [H, P, CNOT, X, Z])
        jobs.append(Job(n, d, ir, shots=shots, seed=seed, stream=c))
    return Workload("uneven_batch_4096", jobs,
                    f"{circuits} circuits n~U[{n_lo},{n_hi}] d~U{set(dims)} 10n ops {{H,P,SUM,X,Z}} + M all, {shots} shots")


def large_tableau(seed=4, n=8192, d=5, shots=32, layers=16, meas_per_layer=64):
    """d=5 n=8192: 4n random {H,P,SUM}, then 16 x (n/2 random SUM + 64 single-qudit M), 32 shots."""
    rng = np.random.default_rng(seed)
    body = random_clifford_ir(rng, n, d, 4 * n, [H, P, CNOT], measure_rounds=0)
    parts = [body]
    for _ in range(layers):
        parts.append(random_clifford_ir(rng, n, d, n // 2, [CNOT], measure_rounds=0))
        qs = rng.choice(n, meas_per_layer, replace=False)
        parts.append(np.stack([np.full(meas_per_layer, M), qs, np.full(meas_per_layer, -1)], axis=1))
    ir = np.concatenate(parts).astype(np.int64)
    return Workload("large_d5_n8192", [Job(n, d, ir, shots=shots, seed=seed, stream=0)],
                    "d=5 n=8192 4n {H,P,SUM} + 16x(n/2 SUM + 64 M), 32 shots")


def published_random(seed=6, gates=100_000, dims=(2, 3, 5, 7), ns=tuple(range(2, 15))):
    """The reference's published timing workload (tests/testdata/simulation_data.json:2-28):
    for every d in {2,3,5,7} and n in 2..14 one random circuit of 100,000 gates drawn
    CNOT 40% / H 30% / P 30%, no measurement, one shot.  Pairs and qudits are drawn as in
    random_clifford_ir."""
    rng = np.random.default_rng(seed)
    jobs = []
    for d in dims:
        for n in ns:
            g = rng.choice(np.array([CNOT, H, P], np.int64), size=gates, p=[0.4, 0.3, 0.3])
            q = rng.integers(n, size=gates)
            t = rng.integers(max(n - 1, 1), size=gates)
            t = np.where(g == CNOT, t + (t >= q), -1)
            jobs.append(Job(n, d, np.stack([g, q, t], axis=1).astype(np.int64), shots=1, seed=seed,
                            stream=len(jobs)))
    return Workload("published_random_1e5", jobs,
                    "d in {2,3,5,7} x n in 2..14: 100,000 gates CNOT 40% / H 30% / P 30%, no measurement, 1 shot")


def serving_request(rng, rid, seed=5, n_lo=8, n_hi=2048, dims=(2, 3, 5, 7, 11), shots_hi=4096):
    """One serving request: n log-uniform in [8, 2048], d in {2,3,5,7,11}, depth 8n, shots log-uniform [1, 4096]."""
    n = int(round(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi)))))
    d = int(rng.choice(dims))
    shots = int(round(np.exp(rng.uniform(0.0, np.log(shots_hi)))))
    ir = random_clifford_ir(rng, n, d, 8 * n, [H, P, CNOT, X, Z])
    return Job(n, d, ir, shots=max(shots, 1), seed=seed, stream=rid)
