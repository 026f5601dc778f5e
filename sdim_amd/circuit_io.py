"""``.chp`` circuit files (sdim/circuit_io.py:6-139).

Format: free-text header, a line containing only ``#``, an optional ``d <dimension>``
line, then one gate per line: ``NAME q [t] [key=value ...]``.  The number of qudits
is one more than the largest index in the file.

``read_circuit`` builds a ``Circuit`` as the reference does; ``read_ir`` goes from the
file straight to the device gate list (the IR triples of ``Program._build_ir``,
sdim/program.py:456-526, plus per-op noise parameters) without creating a
``CircuitInstruction`` per line, and ``read_job`` wraps that in an engine ``Job``.
"""
import os
from dataclasses import dataclass

import numpy as np

from .circuit import CHANNEL_CODES, Circuit
from .gatedata import _DEFAULTS, _REGISTRY, NOISE_GATE

_PKG_PARENT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..")


def _resolve(filename):
    if os.path.exists(filename):
        return filename
    # the reference resolves paths relative to the package's parent directory
    return os.path.join(_PKG_PARENT, filename)


def read_circuit(filename):
    with open(_resolve(filename), "r") as fh:
        lines = fh.readlines()
    start = next(i for i, line in enumerate(lines) if line.strip() == "#")
    body = [line for line in lines[start + 1:]]
    dimension = 2
    first = body[0].split() if body else []
    if first and first[0].upper() == "D":
        dimension = int(first[1])
        body = body[1:]
    max_index = max(int(tok) for line in body for tok in line.split() if tok.isdigit())
    circuit = Circuit(max_index + 1, dimension)
    for line in body:
        if not line.strip():
            continue
        parts = line.split()
        name = parts[0].upper()
        qudits = [int(tok) for tok in parts[1:] if tok.isdigit()]
        extra = [tok for tok in parts[1:] if "=" in tok]
        params = None
        if extra:
            params = {}
            for tok in extra:
                kv = tok.split("=")
                if len(kv) != 2:
                    raise ValueError("Extra parameter doesn't have the correct format.")
                params[kv[0]] = kv[1]
        if len(qudits) == 1:
            circuit.add_gate(name, qudits[0], **(params or {}))
        elif len(qudits) == 2:
            circuit.add_gate(name, qudits[0], qudits[1], **(params or {}))
        else:
            raise ValueError(f"Unexpected number of arguments for gate {name}")
    return circuit


# upper-case name or alias -> (gate id, qudits acted on), as GateData resolves them
_IDS = {}
for _gid, (_name, _args, _aliases) in enumerate(_REGISTRY):
    for _a in (_name,) + _aliases:
        _IDS[_a] = (_gid, _args)


@dataclass
class CircuitIR:
    """A circuit as the device gate list: ``ir`` int64 [ops, 3] (gate id, qudit, target or -1),
    ``noise_channel`` int32 and ``noise_prob`` float64 per op (read for N1 only)."""
    num_qudits: int
    dimension: int
    ir: np.ndarray
    noise_channel: np.ndarray
    noise_prob: np.ndarray


def read_ir(filename):
    """The circuit of a ``.chp`` file as IR arrays, parsed in one pass over the lines with the
    semantics of ``read_circuit`` followed by ``Circuit.add_gate`` and ``_build_ir``
    (sdim/circuit_io.py:6-89, sdim/circuit.py:76-126, sdim/program.py:456-526): gate names and
    aliases case-insensitive, N1's default parameters (channel 'd', prob 0.01) unless given,
    keyword parameters of two-qudit gates dropped, the same errors."""
    with open(_resolve(filename), "r") as fh:
        lines = fh.readlines()
    start = next(i for i, line in enumerate(lines) if line.strip() == "#")
    body = lines[start + 1:]
    dimension = 2
    first = body[0].split() if body else []
    if first and first[0].upper() == "D":
        dimension = int(first[1])
        body = body[1:]
    gid = np.zeros(len(body), np.int64)
    q = np.zeros(len(body), np.int64)
    t = np.full(len(body), -1, np.int64)
    chan = np.zeros(len(body), np.int32)
    prob = np.zeros(len(body), np.float64)
    k = 0
    max_index = -1
    for line in body:
        parts = line.split()
        for tok in parts:
            if tok.isdigit():
                max_index = max(max_index, int(tok))
        if not parts:
            continue
        name = parts[0].upper()
        qudits = [int(tok) for tok in parts[1:] if tok.isdigit()]
        extra = [tok for tok in parts[1:] if "=" in tok]
        params = {}
        for tok in extra:
            kv = tok.split("=")
            if len(kv) != 2:
                raise ValueError("Extra parameter doesn't have the correct format.")
            params[kv[0]] = kv[1]
        if len(qudits) not in (1, 2):
            raise ValueError(f"Unexpected number of arguments for gate {name}")
        if name not in _IDS:
            raise ValueError(f"Gate {name} not found")
        g = _IDS[name][0]
        gid[k] = g
        q[k] = qudits[0]
        if len(qudits) == 2:
            t[k] = qudits[1]
            if g == NOISE_GATE:      # two-qudit calls drop parameters: the IR's defaults apply
                chan[k], prob[k] = CHANNEL_CODES["d"], 0.01
        elif g == NOISE_GATE:
            # the noise gate's registered defaults fill what the line does not give
            p = dict(_DEFAULTS["N1"])
            p.update(params)
            ch = p.get("noise_channel", p.get("channel", "d"))
            if ch not in CHANNEL_CODES:
                raise ValueError("Must specify a valid noise channel.")
            chan[k] = CHANNEL_CODES[ch]
            prob[k] = float(p.get("prob", 0.01))
        k += 1
    if max_index < 0:
        raise ValueError("max() arg is an empty sequence")
    ir = np.stack([gid[:k], q[:k], t[:k]], axis=1)
    return CircuitIR(max_index + 1, dimension, ir, chan[:k], prob[:k])


def read_job(filename, shots=1, seed=0, stream=0):
    """An engine ``Job`` of the circuit in a ``.chp`` file (see ``read_ir``)."""
    from .engine import Job
    c = read_ir(filename)
    if c.num_qudits < 1:
        raise ValueError("Number of qudits must be greater than 0")
    if c.dimension < 2:
        raise ValueError("Dimension must be greater than 1")
    return Job(c.num_qudits, c.dimension, c.ir, shots=shots, seed=seed, stream=stream,
               noise_channel=c.noise_channel, noise_prob=c.noise_prob)


def write_circuit(circuit, output_file="random_circuit.chp", comment="", directory=None):
    text = (f"{comment}\n#\n" if comment else "Randomly-generated Clifford group quantum circuit\n#\n")
    text += f"d {circuit.dimension}\n"
    for g in circuit.operations:
        line = g.gate_name
        line += f" {g.qudit_index} {g.target_index}" if g.target_index is not None else f" {g.qudit_index}"
        if g.params is not None:
            for key, value in g.params.items():
                line += f" {key}={value}"
        text += line + "\n"
    if directory is None:
        directory = os.path.join(_PKG_PARENT, "circuits")
    os.makedirs(directory, exist_ok=True)
    path = os.path.join(directory, output_file)
    with open(path, "w") as fh:
        fh.write(text)
    return path
