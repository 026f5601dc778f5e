"""Pins of the Pauli-frame oracle (oracle/frame.py and oracle_frame in
oracle/sdim_oracle.c), the restatement of sdim/program.py:45-165.

* The two restatements agree record for record (all gates, M_X, RESET, N1).
* Physics fixes the frame sampler per shot on circuits whose measurements are all
  deterministic: a shot's M / M_X values equal the tableau path's values for the
  same shot and the same N1 draws (the tableau oracle is pinned separately), so a
  wrong sign or operand in any conjugation rule, or a wrong noise exponent, fails.
  RESET records repeat the reference record (program.py:150-161).
* Random outcomes: on tiny noiseless circuits the frame shots' M / M_X tuples cover
  exactly the support of the exact statevector distribution, uniformly (stabilizer
  outcome distributions are uniform on an affine subspace), so a missing or wrong
  z randomization fails.
* The reference's statistical tests of the frame path (tests/test_noise_and_io.py,
  tests/test_reset.py) at their tolerance.
"""
import numpy as np
import pytest

from oracle import c_oracle, philox
from oracle import frame as F
from oracle import statevector as SV

INV = {1: 2, 2: 1, 3: 4, 4: 3, 5: 6, 6: 5, 7: 8, 8: 7, 9: 10, 10: 9, 11: 12, 12: 11, 13: 13}


def _noise(ir, ch, pr):
    return {k: (int(ch[k]), philox.keep_threshold53(float(pr[k]))) for k in range(len(ir)) if ir[k][0] == 17}


def _reference(n, d, ir, seed, stream=0):
    """The reference shot: one noiseless tableau shot (N1 acts as I, program.py:31)."""
    ir0 = np.array(ir, dtype=np.int64).reshape(-1, 3).copy()
    ir0[ir0[:, 0] == 17, 0] = 0
    rec, _ = c_oracle.simulate(n, d, ir0, shots=1, seed=seed, stream=stream)
    return [(int(r[4]), bool(r[3])) for r in rec[0]]


def _random_ops(rng, n, d, depth):
    ops, ch, pr = [], [], []
    for _ in range(depth):
        r = rng.random()
        if r < 0.15:
            ops.append((int(rng.choice([14, 15, 16])), int(rng.integers(n)), -1))
        elif r < 0.3:
            ops.append((17, int(rng.integers(n)), -1))
        elif r < 0.65 or n < 2:
            ops.append((int(rng.integers(0, 9)), int(rng.integers(n)), -1))
        else:
            a, b = rng.choice(n, 2, replace=False)
            ops.append((int(rng.integers(9, 14)), int(a), int(b)))
        ch.append(int(rng.integers(3)))
        pr.append(float(rng.random()))
    return np.array(ops, np.int64).reshape(-1, 3), np.array(ch, np.int32), np.array(pr)


@pytest.mark.parametrize("d", [2, 3, 5, 7, 11])
def test_python_and_c_frame_oracles_agree(d):
    rng = np.random.default_rng(700 + d)
    for _ in range(4):
        n = int(rng.integers(1, 7))
        ir, ch, pr = _random_ops(rng, n, d, 60)
        seed = int(rng.integers(1 << 62))
        ref = _reference(n, d, ir, seed, 3)
        nz = _noise(ir, ch, pr)
        a = F.simulate_frame(n, d, ir, ref, 6, seed, 3, nz, shot0=1)
        b = c_oracle.frame(n, d, ir, ref, 6, seed, 3, nz, shot0=1)
        assert np.array_equal(a, b)
        assert np.array_equal(a[:, :, 2], np.broadcast_to(np.arange(1, 7)[:, None], a.shape[:2]))


def _deterministic_circuit(rng, n, depth, p, blocks=2):
    """blocks x [U, U^-1 (N1 after every gate), measure or reset every qudit]: every
    measurement of the noiseless circuit is deterministic."""
    ops, ch, pr = [], [], []

    def add(op, noisy):
        ops.append(op); ch.append(0); pr.append(0.0)
        if noisy:
            for q in ((op[1], op[2]) if op[2] >= 0 else (op[1],)):
                ops.append((17, q, -1)); ch.append(int(rng.integers(3))); pr.append(p)

    for _ in range(blocks):
        fw = []
        for _ in range(depth):
            g = int(rng.integers(1, 14))
            if g >= 9:
                a, b = rng.choice(n, 2, replace=False)
                fw.append((g, int(a), int(b)))
            else:
                fw.append((g, int(rng.integers(n)), -1))
        for op in fw:
            add(op, True)
        for op in reversed(fw):
            add((INV[op[0]], op[1], op[2]), True)
        for q in range(n):
            add((int(rng.choice([14, 16])) if q else 14, q, -1), False)
    return np.array(ops, np.int64), np.array(ch, np.int32), np.array(pr)


@pytest.mark.parametrize("d", [2, 3, 5, 7, 11])
def test_frame_equals_tableau_per_shot_when_measurements_are_deterministic(d):
    rng = np.random.default_rng(800 + d)
    for _ in range(4):
        n = int(rng.integers(2, 6))
        ir, ch, pr = _deterministic_circuit(rng, n, 25, 0.25)
        seed = int(rng.integers(1 << 62))
        ref = _reference(n, d, ir, seed)
        assert all(det for _, det in ref)
        nz = _noise(ir, ch, pr)
        S = 48
        tab, _ = c_oracle.simulate(n, d, ir, shots=S, seed=seed, stream=0, noise=nz)
        fr = c_oracle.frame(n, d, ir, ref, S - 1, seed, 0, nz, shot0=1)
        gid = ir[np.isin(ir[:, 0], (14, 15, 16)), 0]
        meas = gid != 16
        assert np.array_equal(fr[:, meas], tab[1:, meas])
        # the noise actually moved some outcomes
        assert (tab[1:, meas, 4] != np.array([v for v, _ in ref])[meas]).any()
        refv = np.array([v for v, _ in ref])
        assert np.array_equal(fr[:, ~meas, 4], np.broadcast_to(refv[~meas], fr[:, ~meas, 4].shape))


def _tiny_clifford(rng, n, depth):
    ops = []
    for _ in range(depth):
        r = rng.random()
        if r < 0.2:
            ops.append((int(rng.choice([14, 15, 16])), int(rng.integers(n)), -1))
        elif r < 0.6 or n < 2:
            ops.append((int(rng.integers(1, 9)), int(rng.integers(n)), -1))
        else:
            a, b = rng.choice(n, 2, replace=False)
            ops.append((int(rng.integers(9, 14)), int(a), int(b)))
    for q in range(n):
        ops.append((14, q, -1))
    return np.array(ops, np.int64)


@pytest.mark.parametrize("d", [2, 3, 5])
def test_frame_support_matches_statevector(d):
    rng = np.random.default_rng(900 + d)
    for trial in range(5):
        n = 2 if d == 5 else 3
        ir = _tiny_clifford(rng, n, 10)
        gid = ir[np.isin(ir[:, 0], (14, 15, 16)), 0]
        meas = np.flatnonzero(gid != 16)
        dist = SV.outcome_distribution(n, d, ir)
        marg = {}
        for k, p in dist.items():
            key = tuple(k[i] for i in meas)
            marg[key] = marg.get(key, 0.0) + p
        support = {k for k, p in marg.items() if p > 1e-9}
        seed = int(rng.integers(1 << 62))
        ref = _reference(n, d, ir, seed)
        S = 400 * len(support)
        fr = c_oracle.frame(n, d, ir, ref, S, seed, 0, {}, shot0=1)
        seen = {}
        for row in fr[:, meas, 4]:
            key = tuple(int(v) for v in row)
            seen[key] = seen.get(key, 0) + 1
        assert set(seen) == support, f"trial {trial}: frame outcomes {sorted(seen)} vs support {sorted(support)}"
        # uniform on the support, within 6 standard deviations
        exp = S / len(support)
        sd = np.sqrt(S * (1 / len(support)) * (1 - 1 / len(support)))
        for key, c in seen.items():
            assert abs(c - exp) <= 6 * sd + 1, (trial, key, c, exp)


def _counts(vals, d):
    return np.bincount(np.asarray(vals), minlength=d)


@pytest.mark.parametrize("d", [3, 5, 7, 11])
def test_reference_statistics_single_errors(d):
    """tests/test_noise_and_io.py:104-143 and :167-188 restated on the frame oracle
    (100000 shots, |empirical - ideal| < 0.01)."""
    rng = np.random.default_rng(1000 + d)
    shots = 100000
    for channel, code in (("f", 1), ("p", 2)):
        p = float(rng.uniform(0, 1))
        ops = [(5, 0, -1), (17, 0, -1), (6, 0, -1), (14, 0, -1)] if channel == "p" else [(17, 0, -1), (14, 0, -1)]
        ir = np.array(ops, np.int64)
        k = int(np.flatnonzero(ir[:, 0] == 17)[0])
        nz = {k: (code, philox.keep_threshold53(p))}
        seed = int(rng.integers(1 << 62))
        ref = _reference(1, d, ir, seed)
        fr = c_oracle.frame(1, d, ir, ref, shots - 1, seed, 0, nz, shot0=1)
        vals = np.concatenate([[ref[0][0]], fr[:, 0, 4]])
        emp = _counts(vals, d) / shots
        ideal = np.array([1 - p] + [p / (d - 1)] * (d - 1))
        assert np.all(np.abs(emp - ideal) < 0.01), (channel, p, emp, ideal)
    # depolarizing
    p = float(rng.uniform(0, (d * d - 1) / (d * d)))
    ir = np.array([(17, 0, -1), (14, 0, -1)], np.int64)
    nz = {0: (0, philox.keep_threshold53(p))}
    seed = int(rng.integers(1 << 62))
    ref = _reference(1, d, ir, seed)
    fr = c_oracle.frame(1, d, ir, ref, shots - 1, seed, 0, nz, shot0=1)
    emp = _counts(np.concatenate([[ref[0][0]], fr[:, 0, 4]]), d) / shots
    q = p / (d * d - 1)
    ideal = np.array([(1 - p) + (d - 1) * q] + [d * q] * (d - 1))
    assert np.all(np.abs(emp - ideal) < 0.01)


@pytest.mark.parametrize("d", [3, 5, 7, 11])
def test_reference_reset_after_maximal_noise(d):
    """tests/test_reset.py:11-38 restated: maximal depolarizing noise, RESET, X^k, M -> k."""
    rng = np.random.default_rng(1100 + d)
    num_x = int(rng.integers(d))
    ops = [(17, 0, -1), (16, 0, -1)] + [(1, 0, -1)] * num_x + [(14, 0, -1)]
    ir = np.array(ops, np.int64)
    nz = {0: (0, philox.keep_threshold53((d * d - 1) / (d * d)))}
    seed = int(rng.integers(1 << 62))
    ref = _reference(1, d, ir, seed)
    fr = c_oracle.frame(1, d, ir, ref, 20000, seed, 0, nz, shot0=1)
    assert (fr[:, 1, 4] == num_x).all()
    assert (fr[:, 0, 4] == ref[0][0]).all()
