"""Host-side API of the drop-in (no GPU needed): gate registry, circuit building,
IR and record formats, .chp I/O, the random circuit generator, tableau types,
and the C ABI library's exports.  Expected values follow the reference's own
tests (tests/test_program.py:50-148) and sources (sdim/gatedata.py,
sdim/circuit.py, sdim/program.py)."""
import ctypes
import os
import re

import numpy as np
import pytest

import sdim_amd as S
from sdim_amd import _lib
from sdim_amd.engine import MEASUREMENT_DTYPE

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_gate_ids_and_aliases():
    g = S.GateData(3)
    names = ["I", "X", "X_INV", "Z", "Z_INV", "H", "H_INV", "P", "P_INV", "CNOT", "CNOT_INV", "CZ",
             "CZ_INV", "SWAP", "M", "M_X", "RESET", "N1"]
    assert [g.get_gate_id(n) for n in names] == list(range(18))
    for alias, target in [("R", 5), ("DFT", 5), ("H_DAG", 6), ("S", 7), ("PHASE_INV", 8), ("SUM", 9), ("CX", 9),
                          ("C_INV", 10), ("MEASURE", 14), ("MZ", 14), ("MX", 15), ("MR", 16), ("NOISE1", 17)]:
        assert g.get_gate_id(alias) == target
    assert g.get_gate_id("NOPE") is None
    assert g.get_gate_name(9) == "CNOT"
    assert g.gateMap["N1"].defaults == {"channel": "d", "prob": 0.01}
    with pytest.raises(ValueError):
        g.get_gate_name(99)


def test_circuit_broadcasting_and_errors():
    c = S.Circuit(4, 3)
    c.add_gate("h", 0)
    c.add_gate("CNOT", 0, [1, 2, 3])
    c.add_gate("CX", [1, 2], 3)
    c.add_gate("SUM", [0, 1], [2, 3])
    c.add_gate("M", [0, 1])
    got = [(o.gate_id, o.qudit_index, o.target_index) for o in c.operations]
    assert got == [(5, 0, None), (9, 0, 1), (9, 0, 2), (9, 0, 3), (9, 1, 3), (9, 2, 3), (9, 0, 2), (9, 1, 3),
                   (14, 0, None), (14, 1, None)]
    with pytest.raises(ValueError):
        c.add_gate("CNOT", [0, 1, 2], [1, 2])
    with pytest.raises(ValueError):
        c.add_gate("FOO", 0)
    with pytest.raises(ValueError):
        S.Circuit(0, 2)
    with pytest.raises(ValueError):
        S.Circuit(2, 1)


def test_circuit_reference_quirks():
    c = S.Circuit(2, 3)
    c.add_gate("X", 0)
    d = c * 3                      # repeats in place (sdim/circuit.py:128-141)
    assert d is c and len(c.operations) == 3
    c.add_gate("N1", 0, prob=0.5, noise_channel="f")
    assert c.operations[-1].params == {"prob": 0.5, "noise_channel": "f", "channel": "d"}
    c.add_gate("CNOT", 0, 1, foo=1)  # two-qudit instructions drop keyword parameters (circuit.py:125)
    assert c.operations[-1].params is None
    e = S.Circuit(3, 3) + c
    assert e.num_qudits == 3 and len(e.operations) == len(c.operations)
    with pytest.raises(ValueError):
        S.Circuit(2, 2) + c


def test_build_ir_matches_reference_test():
    # tests/test_program.py:112-148
    c = S.Circuit(dimension=3, num_qudits=2)
    c.add_gate("H", 0)
    c.add_gate("CNOT", 0, 1)
    c.add_gate("M", [0, 1])
    p = S.Program(c)
    ir, noise = p._build_ir(p.circuits, 1)
    assert [tuple(r) for r in ir.tolist()] == [(5, 0, -1), (9, 0, 1), (14, 0, -1), (14, 1, -1)]
    c.add_gate("N1", 0, prob=1.0, noise_channel='f')
    ir, noise = p._build_ir(p.circuits, 1)
    assert [tuple(r) for r in ir.tolist()][-1] == (17, 0, -1)
    assert np.any(noise[0][0][0])
    c.add_gate("N1", 1, prob=1.0, noise_channel='p')
    ir, noise = p._build_ir(p.circuits, 1)
    assert np.any(noise[1][0][1])
    c.add_gate("N1", 1, prob=1.0, noise_channel='d')
    ir, noise = p._build_ir(p.circuits, 1)
    assert len(ir) == 7 and np.any(noise[2][0])


def test_results_to_array_and_combine():
    # tests/test_program.py:50-110
    M = S.MeasurementResult
    res = [[M(0, True, 1), M(0, True, 2)], [M(2, True, 3), M(2, True, 4)]]
    exp = np.array([[(0, 0, 0, True, 1), (0, 1, 0, True, 2)], [(2, 0, 0, True, 3), (2, 1, 0, True, 4)]],
                   dtype=MEASUREMENT_DTYPE)
    np.testing.assert_array_equal(S.Program._results_to_array(res), exp)
    nested = [[[M(0, True, 1)]], [[M(1, True, 2)]]]
    exp = np.array([[(0, 0, 0, True, 1)], [(1, 0, 0, True, 2)]], dtype=MEASUREMENT_DTYPE)
    np.testing.assert_array_equal(S.Program._results_to_array(nested), exp)
    with pytest.raises(ValueError):
        S.Program._results_to_array([])
    c = S.Circuit(dimension=3, num_qudits=2)
    p = S.Program(c)
    p.measurement_results = [[[M(0, True, 0)]], [[M(1, True, 0)]]]
    extra = np.array([[[(0, 0, 0, True, 0)]], [[(1, 0, 0, True, 0)]]], dtype=MEASUREMENT_DTYPE)
    p._combine_results(extra)
    assert p.measurement_results == [[[M(0, True, 0), M(0, True, 0)]], [[M(1, True, 0), M(1, True, 0)]]]


def test_ir_keeps_identity_and_noise_params():
    c = S.Circuit(2, 5)
    c.add_gate("I", 0)
    c.add_gate("N1", 1, prob=0.25, noise_channel="p")
    c.add_gate("N1", 0)                              # defaults: channel 'd', prob 0.01
    ir, ch, pr = c.to_ir()
    assert ir.tolist() == [[0, 0, -1], [17, 1, -1], [17, 0, -1]]
    assert ch.tolist() == [0, 2, 0] and pr.tolist() == [0.0, 0.25, 0.01]
    c.add_gate("N1", 0, noise_channel="q")
    with pytest.raises(ValueError):
        c.to_ir()


def test_chp_roundtrip(tmp_path):
    c = S.Circuit(6, 5)
    c.add_gate("H", 0)
    c.add_gate("CNOT", 0, 5)
    c.add_gate("N1", 2, prob=0.3, noise_channel="f")
    c.add_gate("M", [0, 5])
    path = S.write_circuit(c, "rt.chp", comment="round trip", directory=str(tmp_path))
    r = S.read_circuit(path)
    assert r.dimension == 5 and r.num_qudits == 6
    assert [(o.gate_name, o.qudit_index, o.target_index, o.gate_id) for o in r.operations] == \
           [(o.gate_name, o.qudit_index, o.target_index, o.gate_id) for o in c.operations]
    assert {k: str(v) for k, v in r.operations[2].params.items()} == {k: str(v) for k, v in c.operations[2].params.items()}


def test_reference_sample_circuits_parse():
    epr = S.read_circuit(os.path.join(ROOT, "tests", "golden", "epr.chp"))
    assert epr.num_qudits == 2 and epr.dimension == 2
    assert [(o.gate_id, o.qudit_index, o.target_index) for o in epr.operations] == [(5, 0, None), (9, 0, 1), (14, 1, None)]


def test_random_circuit_generator_is_seeded():
    a = S.generate_random_clifford_circuit(5, 40, 3, measurement_rounds=1, seed=11)
    b = S.generate_random_clifford_circuit(5, 40, 3, measurement_rounds=1, seed=11)
    assert str(a) == str(b) and len(a.operations) == 45
    for o in a.operations:
        if o.gate_id in (9, 10, 11, 12):
            assert o.qudit_index != o.target_index


def test_tableau_types():
    t = S.ExtendedTableau(3, 5)
    assert t.order == 5 and t.phase_order == 1 and not t.even and t.prime
    assert t.is_initial_state()
    assert t.tableau.shape == (7, 6)
    t2 = S.ExtendedTableau(2, 2)
    assert t2.order == 4 and t2.phase_order == 2 and t2.coprime_order == {1, 3}
    t.x_block[0, 0] = 7
    t.modulo()
    assert t.x_block[0, 0] == 2 and not t.is_initial_state()
    # exponentiate: stabilizer column scaled, phase per tableau_prime.py:365-380
    u = S.ExtendedTableau(1, 3)
    u.x_block[0, 0], u.z_block[0, 0], u.phase_vector[0] = 1, 1, 1
    u.exponentiate(0, 2)
    assert (u.x_block[0, 0], u.z_block[0, 0], u.phase_vector[0]) == (2, 2, (1 * 2 + 1 * 2 * 1 // 2) % 3)
    m = S.MeasurementResult(1, False, 2)
    assert str(m) == "Measured qudit (1) as (2) and was random"
    assert m == S.MeasurementResult(1, False, 2)
    with pytest.raises(ValueError):
        m.get_tableau()


def test_composite_dimension_rejected():
    with pytest.raises(NotImplementedError):
        S.Program(S.Circuit(2, 4))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "sdim_hip.h")).read()
    return sorted(set(re.findall(r"\b(sdim_[a-z_]+)\s*\(", text)))


def test_c_abi_exports_every_declared_symbol():
    path = _lib.LIB_PATH
    if not os.path.exists(path):
        pytest.skip("libsdim_hip.so not built")
    lib = ctypes.CDLL(path)
    declared = _declared_symbols()
    assert declared, "no declarations found"
    for name in declared:
        assert hasattr(lib, name), f"{name} declared in include/sdim_hip.h but not exported"
    assert set(declared) == set(_lib.EXPORTS)
    # loads through the package binding as well
    L = _lib.load()
    assert L.sdim_last_error() is not None


def test_autotune_signature_and_cache(tmp_path):
    from sdim_amd import autotune
    from sdim_amd.engine import Job
    rng = np.random.default_rng(0)
    jobs = [Job(5, 3, rng.integers(5, 8, size=(10, 3)).astype(np.int64) * [1, 0, 0] + [0, 1, -1], shots=4)]
    other = [Job(6, 3, jobs[0].ir, shots=4)]
    assert autotune.signature(jobs) == autotune.signature(list(jobs))
    assert autotune.signature(jobs) != autotune.signature(other)
    path = str(tmp_path / "t.json")
    assert autotune.cached(jobs, path) is None
    with open(path, "w") as f:
        f.write('{"%s": {"options": {"hbm_threads": 256}, "ms": 1.0, "default_ms": 2.0}}' % autotune.signature(jobs))
    assert autotune.cached(jobs, path) == {"hbm_threads": 256}

    class FakeEngine:
        def __init__(self):
            self.set = {}

        def set_option(self, k, v):
            self.set[k] = v
    e = FakeEngine()
    assert autotune.apply(e, jobs, path) == {"hbm_threads": 256} and e.set == {"hbm_threads": 256}
    keys = set()
    for k, values in autotune.SPACE:
        ks = k if isinstance(k, tuple) else (k,)
        keys.update(ks)
        for v in values:           # joint coordinates list one value per option
            assert len(v if isinstance(k, tuple) else (v,)) == len(ks)
    assert set(autotune.DEFAULTS) == keys


def _random_chp_text(rng, n, d, ops):
    """A .chp file exercising every gate name and alias (mixed case), N1 with and without
    parameters, and two-index single-qudit lines."""
    names = ["I", "X", "x_inv", "Z", "Z_INV", "h", "R", "DFT", "H_INV", "h_dag", "P", "S", "phase", "P_INV",
             "S_INV", "M", "mz", "M_X", "mx", "RESET", "MR", "N1", "noise1"]
    two = ["CNOT", "sum", "CX", "c", "CNOT_INV", "SUM_INV", "CZ", "cz_inv", "SWAP"]
    lines = ["random file", "#", f"d {d}"]
    for _ in range(ops):
        if rng.random() < 0.35:
            a, b = rng.choice(n, 2, replace=False)
            lines.append(f"{rng.choice(two)} {a} {b}")
        else:
            name = str(rng.choice(names))
            line = f"{name} {rng.integers(n)}"
            if name.upper() in ("N1", "NOISE1") and rng.random() < 0.7:
                line += f" prob={rng.random():.3f}"
                if rng.random() < 0.7:
                    line += f" {rng.choice(['noise_channel', 'channel'])}={rng.choice(['d', 'f', 'p'])}"
            elif rng.random() < 0.05:
                line += f" {rng.integers(n)}"
            lines.append(line)
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("seed", range(6))
def test_chp_straight_to_ir_equals_circuit_path(tmp_path, seed):
    """read_ir (file -> IR arrays, no CircuitInstruction objects) against read_circuit +
    add_gate + the IR builder on generated files and the reference's sample circuits."""
    from sdim_amd.circuit import ops_to_ir
    from sdim_amd.circuit_io import read_ir
    rng = np.random.default_rng(seed)
    files = [os.path.join(ROOT, "tests", "golden", "epr.chp"), os.path.join(ROOT, "tests", "golden", "css_steane_final.chp")]
    for k in range(4):
        p = tmp_path / f"g{k}.chp"
        p.write_text(_random_chp_text(rng, int(rng.integers(2, 30)), int(rng.choice([2, 3, 5, 7, 11])), 300))
        files.append(str(p))
    for f in files:
        fast = read_ir(f)
        c = S.read_circuit(f)
        ir, ch, pr = ops_to_ir(c.operations)
        assert (fast.num_qudits, fast.dimension) == (c.num_qudits, c.dimension)
        assert np.array_equal(fast.ir, ir) and np.array_equal(fast.noise_channel, ch)
        assert np.array_equal(fast.noise_prob, pr), f


def test_chp_straight_to_ir_errors(tmp_path):
    from sdim_amd.circuit_io import read_ir
    for body, msg in (("FOO 1\n", "Gate FOO not found"), ("H 1 2 3\n", "Unexpected number of arguments"),
                      ("N1 0 prob=0.1 noise_channel=q\n", "valid noise channel"),
                      ("N1 0 prob=0.1=2\n", "correct format")):
        p = tmp_path / "bad.chp"
        p.write_text("bad\n#\n" + body)
        with pytest.raises(ValueError, match=msg):
            read_ir(str(p))
