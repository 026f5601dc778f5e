"""sdim_amd - MI355X (gfx950) engine for sdim's prime-dimension stabilizer tableau path.

Drop-in API (same names and behaviour as the reference package ``sdim``):

    from sdim_amd import Circuit, Program
    c = Circuit(4, 3)
    c.add_gate("H", 0)
    c.add_gate("CNOT", 0, [1, 2, 3])
    c.add_gate("M", [0, 1, 2, 3])
    Program(c).simulate(shots=1000)

Every shot runs on the GPU through the C ABI in include/sdim_hip.h
(libsdim_hip.so, built with ``python -m sdim_amd.build``).
"""
from .circuit import Circuit, CircuitInstruction
from .circuit_io import CircuitIR, read_circuit, read_ir, read_job, write_circuit
from .engine import MEASUREMENT_DTYPE, Engine, Job, Ticket, default_engine
from .gatedata import Gate, GateData
from .program import Program, SimulationOptions
from .random_circuit import generate_and_write_random_circuit, generate_random_clifford_circuit
from .tableau import ExtendedTableau, MeasurementResult, Tableau

__all__ = [
    "Circuit", "CircuitInstruction", "read_circuit", "write_circuit", "read_ir", "read_job", "CircuitIR",
    "MEASUREMENT_DTYPE", "Engine", "Job",
    "Ticket", "default_engine", "Gate", "GateData", "Program", "SimulationOptions",
    "generate_and_write_random_circuit", "generate_random_clifford_circuit", "ExtendedTableau",
    "MeasurementResult", "Tableau",
]
