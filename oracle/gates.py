"""Gate-id dispatch of the oracle: the reference's prime-branch primitive words.

Gate ids follow sdim/gatedata.py:42-102 (insertion order):
  I 0, X 1, X_INV 2, Z 3, Z_INV 4, H 5, H_INV 6, P 7, P_INV 8, CNOT 9,
  CNOT_INV 10, CZ 11, CZ_INV 12, SWAP 13, M 14, M_X 15, RESET 16, N1 17.

Each apply_* restates the corresponding prime branch of
sdim/tableau/tableau_gates.py; the line range is in the docstring.
"""


def apply_X(t, q):
    """tableau_gates.py:27-54 (even: H P P H; odd: H P^-1 H H P H)."""
    if t.even:
        t.hadamard(q); t.phase(q); t.phase(q); t.hadamard(q)
    else:
        t.hadamard(q); t.phase_inv(q); t.hadamard(q); t.hadamard(q); t.phase(q); t.hadamard(q)


def apply_X_inv(t, q):
    """tableau_gates.py:56-83 (even: H P P H; odd: H P H H P^-1 H)."""
    if t.even:
        t.hadamard(q); t.phase(q); t.phase(q); t.hadamard(q)
    else:
        t.hadamard(q); t.phase(q); t.hadamard(q); t.hadamard(q); t.phase_inv(q); t.hadamard(q)


def apply_Z(t, q):
    """tableau_gates.py:85-110 (even: P P; odd: P^-1 H H P H H)."""
    if t.even:
        t.phase(q); t.phase(q)
    else:
        t.phase_inv(q); t.hadamard(q); t.hadamard(q); t.phase(q); t.hadamard(q); t.hadamard(q)


def apply_Z_inv(t, q):
    """tableau_gates.py:112-137 (even: P^-1 P^-1; odd: P H^-1 H^-1 P^-1 H^-1 H^-1)."""
    if t.even:
        t.phase_inv(q); t.phase_inv(q)
    else:
        t.phase(q); t.hadamard_inv(q); t.hadamard_inv(q); t.phase_inv(q)
        t.hadamard_inv(q); t.hadamard_inv(q)


def apply_CZ(t, c, tg):
    """tableau_gates.py:229-244: H^-1(t) CNOT(c,t) H(t)."""
    t.hadamard_inv(tg); t.cnot(c, tg); t.hadamard(tg)


def apply_CZ_inv(t, c, tg):
    """tableau_gates.py:246-261: H^-1(t) CNOT^-1(c,t) H(t)."""
    t.hadamard_inv(tg); t.cnot_inv(c, tg); t.hadamard(tg)


def apply_SWAP(t, a, b):
    """tableau_gates.py:319-328 (prime branch, Farinholt Fig. 1)."""
    t.cnot(a, b); t.hadamard(a); t.hadamard(b)
    t.cnot(a, b); t.hadamard(a); t.hadamard(b)
    t.cnot(a, b); t.hadamard(b); t.hadamard(b)


UNITARY_GATES = {
    1: lambda t, q, _: apply_X(t, q),
    2: lambda t, q, _: apply_X_inv(t, q),
    3: lambda t, q, _: apply_Z(t, q),
    4: lambda t, q, _: apply_Z_inv(t, q),
    5: lambda t, q, _: t.hadamard(q),
    6: lambda t, q, _: t.hadamard_inv(q),
    7: lambda t, q, _: t.phase(q),
    8: lambda t, q, _: t.phase_inv(q),
    9: lambda t, q, tg: t.cnot(q, tg),
    10: lambda t, q, tg: t.cnot_inv(q, tg),
    11: lambda t, q, tg: apply_CZ(t, q, tg),
    12: lambda t, q, tg: apply_CZ_inv(t, q, tg),
    13: lambda t, q, tg: apply_SWAP(t, q, tg),
}

MEASURE, MEASURE_X, RESET, NOISE = 14, 15, 16, 17
