"""``.chp`` circuit files (sdim/circuit_io.py:6-139).

Format: free-text header, a line containing only ``#``, an optional ``d <dimension>``
line, then one gate per line: ``NAME q [t] [key=value ...]``.  The number of qudits
is one more than the largest index in the file.
"""
import os

from .circuit import Circuit

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
