"""Circuit construction API, matching sdim's ``Circuit`` / ``CircuitInstruction``
(sdim/circuit.py:5-242).

Kept reference behaviours:
  * ``add_gate`` broadcasts one control over many targets, many controls over
    one target, or pairs equal-length lists (sdim/circuit.py:95-125);
  * two-qudit ``add_gate`` calls drop keyword parameters (circuit.py:125);
  * ``Circuit.__mul__`` repeats the operations in place and returns ``self``
    (circuit.py:128-141).
"""
from dataclasses import dataclass
from typing import List, Optional, Union

import numpy as np

from .gatedata import GateData


@dataclass
class CircuitInstruction:
    gate_data: GateData
    gate_name: str
    qudit_index: int
    target_index: int = None
    gate_id: int = None
    name: str = None
    params: Optional[dict] = None

    def __post_init__(self):
        self.gate_id = self.gate_data.get_gate_id(self.gate_name)
        if self.gate_id is None:
            raise ValueError(f"Gate {self.gate_name} not found")
        self.name = self.gate_data.get_gate_name(self.gate_id)

    def __str__(self):
        return f"{self.gate_id} {self.qudit_index} {self.target_index}"


@dataclass
class Circuit:
    num_qudits: int
    dimension: int = 2
    operations: list = None
    gate_data: GateData = None

    def __post_init__(self):
        if self.num_qudits < 1:
            raise ValueError("Number of qudits must be greater than 0")
        if self.dimension < 2:
            raise ValueError("Dimension must be greater than 1")
        self.operations = self.operations or []
        self.gate_data = self.gate_data or GateData(self.dimension)

    def add_gate(self, gate_name: str, control: Union[int, List[int]],
                 target: Union[int, List[int], None] = None, **kwargs):
        controls = [control] if isinstance(control, (int, np.integer)) else list(control)
        targets = [target] if isinstance(target, (int, np.integer)) else target
        upper = gate_name.upper()
        canonical = self.gate_data.aliasMap.get(upper, upper)
        gate = self.gate_data.gateMap.get(canonical)
        if gate is not None and gate.defaults:
            for key, value in gate.defaults.items():
                kwargs.setdefault(key, value)
        if targets is None:
            for c in controls:
                self.operations.append(CircuitInstruction(self.gate_data, upper, int(c), None, params=kwargs))
            return
        targets = list(targets)
        if len(controls) == 1:
            pairs = [(controls[0], t) for t in targets]
        elif len(targets) == 1:
            pairs = [(c, targets[0]) for c in controls]
        elif len(controls) == len(targets):
            pairs = list(zip(controls, targets))
        else:
            raise ValueError("Invalid combination of control and target qubits")
        for c, t in pairs:
            self.operations.append(CircuitInstruction(self.gate_data, upper, int(c), int(t)))

    def __mul__(self, repetitions: int):
        original = list(self.operations)
        for _ in range(repetitions - 1):
            self.operations.extend(original)
        return self

    def __imul__(self, repetitions: int):
        return self.__mul__(repetitions)

    def __add__(self, other):
        if self.dimension != other.dimension:
            raise ValueError("Cannot add circuits with different dimensions")
        out = Circuit(max(self.num_qudits, other.num_qudits), self.dimension)
        out.operations = list(self.operations) + list(other.operations)
        return out

    def __iadd__(self, other):
        if self.dimension != other.dimension:
            raise ValueError("Cannot add circuits with different dimensions")
        self.num_qudits = max(self.num_qudits, other.num_qudits)
        self.operations.extend(other.operations)
        return self

    def __str__(self):
        return "\n".join(str(op) for op in self.operations)

    def print_gateData(self):
        print(self.gate_data)

    @classmethod
    def from_operation_list(cls, operation_list, num_qudits, dimension):
        circuit = cls(num_qudits, dimension)
        for op in operation_list:
            if isinstance(op, tuple):
                name, qudits = op[0], op[1]
                if len(qudits) == 1:
                    circuit.add_gate(name, qudits[0])
                elif len(qudits) == 2:
                    circuit.add_gate(name, qudits[0], qudits[1])
                else:
                    raise ValueError(f"Unsupported number of qudits for gate {name}")
            elif isinstance(op, CircuitInstruction):
                circuit.add_gate(op.gate_name, op.qudit_index, op.target_index)
            else:
                raise ValueError(f"Unsupported operation type: {type(op)}")
        return circuit

    # ------------------------------------------------------------ IR
    def to_ir(self):
        """(gate_id, qudit, target | -1) int64 triples plus per-op noise channel / probability.

        Unlike ``Program._build_ir`` the identity gate is kept, so that op indices
        (which key the Philox draws) equal instruction indices.
        """
        return ops_to_ir(self.operations)


CHANNEL_CODES = {"d": 0, "f": 1, "p": 2}


def ops_to_ir(operations):
    k = len(operations)
    ir = np.empty((k, 3), dtype=np.int64)
    chan = np.zeros(k, dtype=np.int32)
    prob = np.zeros(k, dtype=np.float64)
    for i, op in enumerate(operations):
        ir[i, 0] = op.gate_id
        ir[i, 1] = op.qudit_index
        ir[i, 2] = -1 if op.target_index is None else op.target_index
        if op.gate_id == 17:
            params = op.params or {}
            ch = params.get("noise_channel", params.get("channel", "d"))
            if ch not in CHANNEL_CODES:
                raise ValueError("Must specify a valid noise channel.")
            chan[i] = CHANNEL_CODES[ch]
            prob[i] = float(params.get("prob", 0.01))
    return ir, chan, prob
