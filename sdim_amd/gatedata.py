"""Gate registry: names, aliases and the numeric gate ids of the IR.

The ids are the ABI of the device gate list and must equal the reference's
(sdim/gatedata.py:42-102, ids assigned in registration order):

    I 0, X 1, X_INV 2, Z 3, Z_INV 4, H 5, H_INV 6, P 7, P_INV 8, CNOT 9,
    CNOT_INV 10, CZ 11, CZ_INV 12, SWAP 13, M 14, M_X 15, RESET 16, N1 17
"""
from dataclasses import dataclass
from typing import Optional

# (canonical name, qudits acted on, aliases) in id order
_REGISTRY = (
    ("I", 1, ()),
    ("X", 1, ()),
    ("X_INV", 1, ()),
    ("Z", 1, ()),
    ("Z_INV", 1, ()),
    ("H", 1, ("R", "DFT")),
    ("H_INV", 1, ("R_INV", "DFT_INV", "H_DAG", "R_DAG", "DFT_DAG")),
    ("P", 1, ("PHASE", "S")),
    ("P_INV", 1, ("PHASE_INV", "S_INV")),
    ("CNOT", 2, ("SUM", "CX", "C")),
    ("CNOT_INV", 2, ("SUM_INV", "CX_INV", "C_INV")),
    ("CZ", 2, ()),
    ("CZ_INV", 2, ()),
    ("SWAP", 2, ()),
    ("M", 1, ("MEASURE", "COLLAPSE", "MZ")),
    ("M_X", 1, ("MEASURE_X", "MX")),
    ("RESET", 1, ("MR", "MEASURE_RESET", "MEASURE_R")),
    ("N1", 1, ("NOISE1",)),
)

# default parameters of the noise gate (sdim/gatedata.py:99-102)
_DEFAULTS = {"N1": {"channel": "d", "prob": 0.01}}

GATE_IDS = {name: gid for gid, (name, _, _) in enumerate(_REGISTRY)}
MEASUREMENT_GATES = (GATE_IDS["M"], GATE_IDS["M_X"], GATE_IDS["RESET"])
NOISE_GATE = GATE_IDS["N1"]


@dataclass
class Gate:
    name: str
    arg_count: int
    gate_id: int
    defaults: Optional[dict] = None

    def __str__(self):
        return f"{self.name} {self.gate_id}"


class GateData:
    """Named gates of a qudit dimension; ``gateMap`` and ``aliasMap`` as in the reference."""

    def __init__(self, dimension=2):
        self.dimension = dimension
        self.gateMap = {}
        self.aliasMap = {}
        for gid, (name, args, aliases) in enumerate(_REGISTRY):
            defaults = dict(_DEFAULTS[name]) if name in _DEFAULTS else None
            self.gateMap[name] = Gate(name, args, gid, defaults)
            for a in aliases:
                self.aliasMap[a] = name
        self.num_gates = len(_REGISTRY)

    def __str__(self):
        return "\n".join(str(g) for g in self.gateMap.values())

    def add_gate(self, name, arg_count):
        self.gateMap[name] = Gate(name, arg_count, self.num_gates)
        self.num_gates += 1

    def add_gate_alias(self, name, list_alias):
        for a in list_alias:
            self.aliasMap[a] = name

    def get_gate_id(self, gate_name):
        if gate_name in self.gateMap:
            return self.gateMap[gate_name].gate_id
        if gate_name in self.aliasMap:
            return self.gateMap[self.aliasMap[gate_name]].gate_id
        return None

    def get_gate_name(self, gate_id):
        for name, g in self.gateMap.items():
            if g.gate_id == gate_id:
                return name
        raise ValueError(f"Gate ID {gate_id} not found")
