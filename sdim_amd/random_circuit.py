"""Seeded random Clifford circuits, following ``generate_random_clifford_circuit``
(sdim/random_circuit.py:9-68): a uniform choice from the gate set with Python's
``random`` module, two-qudit gates on ``random.sample(range(n), 2)``, one-qudit
gates on ``random.randint(0, n - 1)``, then ``measurement_rounds`` rounds of M on
every qudit.  The same seed and gate set give the same circuit as the reference
generator, because the same sequence of ``random`` calls is made."""
import random

from .circuit import Circuit

DEFAULT_GATES = ["H", "P", "CNOT", "X", "Z", "H_INV", "P_INV", "CNOT_INV", "X_INV", "Z_INV", "CZ", "CZ_INV"]
TWO_QUDIT = {"CNOT", "CNOT_INV", "CZ", "CZ_INV"}


def generate_random_clifford_circuit(num_qudits, num_gates, dimension, measurement_rounds=0, seed=None,
                                     gate_set=None):
    if seed is not None:
        random.seed(seed)
    gates = gate_set or DEFAULT_GATES
    circuit = Circuit(num_qudits, dimension)
    for _ in range(num_gates):
        g = random.choice(gates)
        if g in TWO_QUDIT:
            a, b = random.sample(range(num_qudits), 2)
            circuit.add_gate(g, a, b)
        else:
            circuit.add_gate(g, random.randint(0, num_qudits - 1))
    for _ in range(measurement_rounds):
        for q in range(num_qudits):
            circuit.add_gate("M", q)
    return circuit


def generate_and_write_random_circuit(num_qudits, num_gates, dimension, measurement_rounds=0,
                                      output_file="random_circuit.chp", seed=None):
    from .circuit_io import write_circuit
    circuit = generate_random_clifford_circuit(num_qudits, num_gates, dimension, measurement_rounds, seed)
    write_circuit(circuit, output_file)
    return circuit
