"""The drop-in API on the GPU: the reference's own expectations through
``Program.simulate`` (tests/test_circuit.py:9-263 prime cases,
tests/test_program.py:14-47, tests/test_noise_and_io.py:174-200,
tests/test_reset.py:11-36), plus record_tableau snapshots, initial tableaux
and ExtendedTableau methods checked against the oracle."""
import random

import numpy as np
import pytest

import sdim_amd as S
from oracle import c_oracle, philox
from oracle import program as OP
from oracle.tableau import OracleTableau

pytestmark = pytest.mark.gpu
M = S.MeasurementResult


def test_phase_kickback():
    c = S.Circuit(2, 2)
    for g, q in [("H", 1), ("P", 1), ("H", 0)]:
        c.add_gate(g, q)
    c.add_gate("CNOT", 0, 1); c.add_gate("H", 1); c.add_gate("CNOT", 0, 1); c.add_gate("H", 1)
    c.add_gate("P", 0); c.add_gate("H", 0); c.add_gate("M", 0); c.add_gate("P", 1); c.add_gate("H", 1); c.add_gate("M", 1)
    assert S.Program(c).simulate() == [M(0, True, 1), M(1, True, 1)]


def test_qubit_and_qutrit_flip():
    c = S.Circuit(2, 2)
    for g in ("H", "P", "P", "H", "M"):
        c.add_gate(g, 0)
    c.add_gate("X", 1); c.add_gate("M", 1)
    assert S.Program(c).simulate() == [M(0, True, 1), M(1, True, 1)]
    c = S.Circuit(2, 3)
    c.add_gate("X", 0); c.add_gate("M", 0)
    assert S.Program(c).simulate() == [M(0, True, 1)]


@pytest.mark.parametrize("d", [2, 3, 5, 7, 11])
def test_swap_computational_basis(d):
    rng = random.Random(d)
    for _ in range(4):
        x0, x1 = rng.randrange(d), rng.randrange(d)
        c = S.Circuit(2, d)
        for _ in range(x0):
            c.add_gate("X", 0)
        for _ in range(x1):
            c.add_gate("X", 1)
        c.add_gate("SWAP", 0, 1); c.add_gate("M", 0); c.add_gate("M", 1)
        assert S.Program(c).simulate() == [M(0, True, x1), M(1, True, x0)]
        c2 = S.Circuit(2, d)
        for _ in range(x0):
            c2.add_gate("X", 0)
        for _ in range(x1):
            c2.add_gate("X", 1)
        c2.add_gate("SWAP", 0, 1); c2.add_gate("SWAP", 0, 1); c2.add_gate("M", 0); c2.add_gate("M", 1)
        assert S.Program(c2).simulate() == [M(0, True, x0), M(1, True, x1)]


@pytest.mark.parametrize("a", range(3))
@pytest.mark.parametrize("b", range(3))
def test_qutrit_swap_in_x_basis(a, b):
    c = S.Circuit(2, 3)
    for _ in range(a):
        c.add_gate("X", 0)
    c.add_gate("DFT", 0)
    for _ in range(b):
        c.add_gate("X", 1)
    c.add_gate("DFT", 1); c.add_gate("SWAP", 0, 1); c.add_gate("MX", 0); c.add_gate("MX", 1)
    assert S.Program(c).simulate() == [M(0, True, b), M(1, True, a)]


@pytest.mark.parametrize("f0,f1", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_qubit_deutsch(f0, f1):
    c = S.Circuit(2, 2)
    c.add_gate("H", 0); c.add_gate("X", 1); c.add_gate("H", 1)
    if f0:
        c.add_gate("CNOT", 0, 1); c.add_gate("X", 1)
    if f1:
        c.add_gate("CNOT", 0, 1)
    c.add_gate("H", 0); c.add_gate("M", 0)
    assert S.Program(c).simulate() == [M(0, True, 0 if f0 == f1 else 1)]


def test_stabilizer_extraction():
    c = S.Circuit(dimension=3, num_qudits=5)
    c.add_gate("X", [2, 3, 4]); c.add_gate("H", 0)
    c.add_gate("CZ", 0, 2); c.add_gate("CZ", 0, 3); c.add_gate("CZ_INV", 0, 4); c.add_gate("H_INV", 0)
    c.add_gate("H", 1); c.add_gate("CZ_INV", 1, 2); c.add_gate("CZ_INV", 1, 3); c.add_gate("CZ", 1, 4)
    c.add_gate("H_INV", 1); c.add_gate("M", 0); c.add_gate("M", 1)
    assert S.Program(c).simulate() == [M(0, True, 1), M(1, True, 2)]
    c = S.Circuit(dimension=3, num_qudits=5)
    c.add_gate("H", [2, 3, 4]); c.add_gate("Z", 2); c.add_gate("H", 0)
    c.add_gate("CX", 0, 2); c.add_gate("CX", 0, 3); c.add_gate("CX_INV", 0, 4); c.add_gate("H_INV", 0)
    c.add_gate("H", 1); c.add_gate("CX_INV", 1, 2); c.add_gate("CX_INV", 1, 3); c.add_gate("CX", 1, 4)
    c.add_gate("H_INV", 1); c.add_gate("M", 0); c.add_gate("M", 1)
    assert S.Program(c).simulate() == [M(0, True, 2), M(1, True, 1)]


def test_deutsch_qutrit():
    c = S.Circuit(2, 3)
    c.add_gate("H", 0); c.add_gate("X", 1); c.add_gate("H", 1); c.add_gate("CNOT", 0, 1)
    c.add_gate("H_INV", 0); c.add_gate("M", 0)
    p = S.Program(c)
    assert p.simulate() == [M(0, True, 2)]


def test_measurement_formats():
    # tests/test_program.py:14-47
    c = S.Circuit(dimension=3, num_qudits=2)
    c.add_gate("M", [0, 1])
    p = S.Program(c)
    assert p.simulate(shots=1) == [M(0, True, 0), M(1, True, 0)]
    assert p.measurement_results == [[[M(0, True, 0)]], [[M(1, True, 0)]]]
    c.add_gate("M", [0, 1])
    p = S.Program(c)
    assert p.simulate(shots=1) == [M(0, True, 0), M(0, True, 0), M(1, True, 0), M(1, True, 0)]
    r = p.simulate(shots=2, force_tableau=True)
    assert r == [[[M(0, True, 0)] * 2, [M(0, True, 0)] * 2], [[M(1, True, 0)] * 2, [M(1, True, 0)] * 2]]
    assert r == p.measurement_results


@pytest.mark.parametrize("d", [3, 5])
def test_identity_rb_all_zero(d):
    # tests/test_noise_and_io.py:174-200
    inv = {"I": "I", "X": "X_INV", "X_INV": "X", "Z": "Z_INV", "Z_INV": "Z", "H": "H_INV", "H_INV": "H",
           "P": "P_INV", "P_INV": "P"}
    rng = random.Random(d)
    gates = [rng.choice(list(inv)) for _ in range(200)]
    c = S.Circuit(dimension=d, num_qudits=1)
    for g in gates + [inv[g] for g in reversed(gates)]:
        c.add_gate(g, 0)
    c.add_gate("M", 0)
    res = S.Program(c).simulate(shots=1000)
    assert all(m.measurement_value == 0 for m in res[0][0])


@pytest.mark.parametrize("d", [3, 5, 7, 11])
def test_reset_after_maximal_noise(d):
    # tests/test_reset.py:11-36
    p = (d * d - 1) / (d * d)
    for num_x in range(d):
        c = S.Circuit(dimension=d, num_qudits=1)
        c.add_gate("N1", 0, prob=p, noise_channel='d')
        c.add_gate("RESET", 0)
        for _ in range(num_x):
            c.add_gate("X", 0)
        c.add_gate("M", 0)
        res = S.Program(c).simulate(shots=2000)
        assert all(m.measurement_value == num_x for m in res[0][1])


def test_noise_channel_statistics():
    # flip channel frequencies (tests/test_noise_and_io.py:101-131), loose statistical check
    d, p, shots = 5, 0.3, 20000
    c = S.Circuit(dimension=d, num_qudits=1)
    c.add_gate("N1", 0, prob=p, noise_channel='f')
    c.add_gate("M", 0)
    res = S.Program(c, seed=123).simulate(shots=shots)
    counts = np.bincount([m.measurement_value for m in res[0][0]], minlength=d) / shots
    ideal = np.array([1 - p] + [p / (d - 1)] * (d - 1))
    assert np.all(np.abs(counts - ideal) < 0.02)


def _oracle_records(prog, shots):
    ir, chan, prob = prog.circuits[0].to_ir()
    noise = {int(k): (int(chan[k]), philox.keep_threshold53(float(prob[k]))) for k in np.nonzero(ir[:, 0] == 17)[0]}
    t = prog.stabilizer_tableau
    return c_oracle.simulate(t.num_qudits, t.dimension, ir, shots=shots, seed=prog.last_seed,
                             stream=prog.last_stream, noise=noise, keep_shots=shots)


@pytest.mark.parametrize("d", [2, 3, 5])
def test_program_matches_oracle_and_final_tableau(d):
    c = S.generate_random_clifford_circuit(6, 80, d, measurement_rounds=1, seed=d)
    c.add_gate("N1", 2, prob=0.5, noise_channel="d")
    c.add_gate("RESET", 2)
    prog = S.Program(c, seed=99)
    res = prog.simulate(shots=5, force_tableau=True)
    rec, tabs = _oracle_records(prog, 5)
    for k in range(rec.shape[1]):
        q, rnd = int(rec[0, k, 0]), int(rec[0, k, 1])
        for s in range(5):
            m = res[q][rnd][s]
            assert (m.qudit_index, m.deterministic, m.measurement_value) == (q, bool(rec[s, k, 3]), int(rec[s, k, 4]))
    # the program keeps the last shot's final tableau (sdim/program.py:308-351)
    got = prog.stabilizer_tableau.arrays()
    for a, b in zip(got, tabs[-1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("d", [2, 3, 5])
def test_program_frame_path(d):
    """shots > 1 without force_tableau (sdim/program.py:244-265): shot 0 is the noiseless
    tableau shot, shots 1.. are frame shots; the final tableau is the reference shot's."""
    c = S.generate_random_clifford_circuit(6, 80, d, measurement_rounds=2, seed=10 + d)
    c.add_gate("N1", 2, prob=0.5, noise_channel="d")
    c.add_gate("M_X", 1)
    c.add_gate("RESET", 2)
    prog = S.Program(c, seed=77)
    res = prog.simulate(shots=300)
    ir, chan, prob = c.to_ir()
    ir0 = ir.copy()
    ir0[ir0[:, 0] == 17, 0] = 0
    rec0, tabs = c_oracle.simulate(6, d, ir0, shots=1, seed=77, stream=0, keep_shots=1)
    ref = [(int(r[4]), bool(r[3])) for r in rec0[0]]
    noise = {int(k): (int(chan[k]), philox.keep_threshold53(float(prob[k]))) for k in np.nonzero(ir[:, 0] == 17)[0]}
    fr = c_oracle.frame(6, d, ir, ref, 299, 77, 0, noise, shot0=1)
    for k in range(rec0.shape[1]):
        q, rnd = int(rec0[0, k, 0]), int(rec0[0, k, 1])
        got = res[q][rnd]
        assert len(got) == 300
        assert (got[0].qudit_index, got[0].deterministic, got[0].measurement_value) == (q, ref[k][1], ref[k][0])
        for s in range(299):
            m = got[s + 1]
            assert (m.qudit_index, m.deterministic, m.measurement_value) == (q, bool(fr[s, k, 3]), int(fr[s, k, 4]))
    for a, b in zip(prog.stabilizer_tableau.arrays(), tabs[0]):
        assert np.array_equal(a, b)


def test_record_tableau_snapshots():
    from oracle import gates as OG
    c = S.Circuit(3, 3)
    c.add_gate("H", 0); c.add_gate("CNOT", 0, 1); c.add_gate("M", 0); c.add_gate("X", 2); c.add_gate("RESET", 2)
    c.add_gate("M", 1)
    prog = S.Program(c, seed=5)
    res = prog.simulate(shots=2, record_tableau=True)
    ir, _, _ = c.to_ir()
    for shot in range(2):
        # the literal oracle, snapshotting right after each measurement (before RESET's fix-up)
        t = OracleTableau(3, 3)
        snaps = []
        for k, (g, q, tg) in enumerate(ir.tolist()):
            if g in (14, 16):
                _, m = t.measure(q, lambda: philox.measurement_outcome(5, 0, k, shot, 3))
                snap = t.copy()
                snap.modulo()
                snaps.append((q, snap))
                if g == 16:
                    for _ in range((-m) % 3):
                        OG.apply_X(t, q)
            else:
                OG.UNITARY_GATES[g](t, q, tg)
        for q, snap in snaps:
            got = res[q][0][shot].get_tableau()
            for a, b in zip(got.arrays(), snap.arrays()):
                assert np.array_equal(a, b)


def test_initial_tableau_and_extended_tableau_methods():
    # an ExtendedTableau evolved with its own (GPU-backed) methods, then used as Program's initial state
    t = S.ExtendedTableau(3, 5)
    t.hadamard(0); t.cnot(0, 1); t.phase(1); t.hadamard_inv(2); t.cnot_inv(2, 0); t.phase_inv(0)
    o = OracleTableau(3, 5)
    o.hadamard(0); o.cnot(0, 1); o.phase(1); o.hadamard_inv(2); o.cnot_inv(2, 0); o.phase_inv(0)
    o.modulo()
    for a, b in zip(t.arrays(), o.arrays()):
        assert np.array_equal(a, b)
    c = S.Circuit(3, 5)
    c.add_gate("M", [0, 1, 2])
    prog = S.Program(c, tableau=t, seed=7)
    with pytest.raises(NotImplementedError):
        prog.simulate(shots=3)                # the frame sampler starts from |0...0>
    res = prog.simulate(shots=3, force_tableau=True)
    ops = [(14, 0, -1), (14, 1, -1), (14, 2, -1)]
    rec, _ = OP.simulate(3, 5, ops, shots=3, seed=7, stream=0, initial=o)
    got = [(q, rnd, s, m.deterministic, m.measurement_value)
           for q in range(3) for rnd in range(1) for s, m in enumerate(res[q][rnd])]
    exp = sorted(rec, key=lambda r: (r[0], r[1], r[2]))
    assert got == [(q, r, s, d, v) for (q, r, s, d, v) in exp]
    mr = t.measure(1)
    assert isinstance(mr, S.MeasurementResult) and mr.qudit_index == 1


def test_kernel_snapshots_equal_prefix_runs():
    """The kernel's tableau snapshots (per measurement event for record_tableau, per instruction
    for verbose) equal the final tableaux of the circuits cut at those points."""
    from snapshot_helpers import check_snapshots
    from sdim_amd.engine import Engine
    eng = Engine()
    print("[snapshots] random circuits, batched and plain", flush=True)
    check_snapshots(eng, np.random.default_rng(808), count=3)
    eng.close()


def test_verbose_prints_every_instruction(capsys):
    c = S.Circuit(2, 3)
    c.add_gate("H", 0); c.add_gate("CNOT", 0, 1); c.add_gate("M", 1); c.add_gate("RESET", 1)
    S.Program(c, seed=3).simulate(shots=1, verbose=True, show_gate=True)
    out = capsys.readouterr().out
    assert out.count("Time step") == 3 and out.count("Final step") == 1 and "Initial state" in out
