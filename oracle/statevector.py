"""Dense statevector / operator model used only to pin the oracle.

Matrix formulas restated from sdim/unitary.py for prime d:
  X|j> = |j+1 mod d>                      (unitary.py:26-37)
  Z = diag(w^j), w = exp(2 pi i / d)      (unitary.py:39-50)
  H[m, k] = w^(m k) / sqrt(d)             (unitary.py:52-70, prime branch)
  P = diag(w^(j(j-1)/2)) for odd d, diag(w^(j^2/2)) for d = 2   (unitary.py:88-110)
  SUM|c, t> = |c, c + t>                  (unitary.py:112-126)
Inverse gates are adjoints.  Qudit 0 is the most significant tensor factor.

A tableau column (x, z, phi) denotes the operator tau^phi * prod_k X_k^x_k Z_k^z_k
with tau = exp(2 pi i / (po d)), po = 2 for d = 2 else 1 (the phase is counted
in units of w^(1/po), dataclasses.py:88-106).
"""
import numpy as np


def omega(d):
    return np.exp(2j * np.pi / d)


def one_qudit(name, d):
    w = omega(d)
    j = np.arange(d)
    if name == "X":
        m = np.zeros((d, d), dtype=complex)
        for i in range(d):
            m[(i + 1) % d, i] = 1
        return m
    if name == "Z":
        return np.diag(w ** j)
    if name == "H":
        return w ** np.outer(j, j) / np.sqrt(d)
    if name == "P":
        if d % 2 == 1:
            return np.diag(w ** (j * (j - 1) / 2))
        return np.diag(w ** (j * j / 2))
    raise KeyError(name)


def sum_gate(d):
    m = np.zeros((d * d, d * d), dtype=complex)
    for c in range(d):
        for t in range(d):
            m[d * c + (c + t) % d, d * c + t] = 1
    return m


GATE_MATRIX = {
    1: lambda d: one_qudit("X", d),
    2: lambda d: one_qudit("X", d).conj().T,
    3: lambda d: one_qudit("Z", d),
    4: lambda d: one_qudit("Z", d).conj().T,
    5: lambda d: one_qudit("H", d),
    6: lambda d: one_qudit("H", d).conj().T,
    7: lambda d: one_qudit("P", d),
    8: lambda d: one_qudit("P", d).conj().T,
    9: lambda d: sum_gate(d),
    10: lambda d: sum_gate(d).conj().T,
}


def cz_matrix(d, inverse=False):
    """H^-1 on the target, SUM (or SUM^-1), then H on the target (tableau_gates.py:229-261)."""
    h = one_qudit("H", d)
    ih = np.kron(np.eye(d), h)
    s = sum_gate(d)
    if inverse:
        s = s.conj().T
    return ih @ s @ ih.conj().T


def apply_1q(psi, u, q, n, d):
    psi = psi.reshape((d,) * n)
    psi = np.moveaxis(psi, q, 0)
    psi = np.tensordot(u, psi, axes=([1], [0]))
    psi = np.moveaxis(psi, 0, q)
    return psi.reshape(-1)


def apply_2q(psi, u, a, b, n, d):
    psi = psi.reshape((d,) * n)
    psi = np.moveaxis(psi, (a, b), (0, 1))
    shp = psi.shape
    psi = (u @ psi.reshape(d * d, -1)).reshape(shp)
    psi = np.moveaxis(psi, (0, 1), (a, b))
    return psi.reshape(-1)


def apply_gate(psi, gid, q, t, n, d):
    if gid in (1, 2, 3, 4, 5, 6, 7, 8):
        return apply_1q(psi, GATE_MATRIX[gid](d), q, n, d)
    if gid in (9, 10):
        return apply_2q(psi, GATE_MATRIX[gid](d), q, t, n, d)
    if gid in (11, 12):
        return apply_2q(psi, cz_matrix(d, inverse=(gid == 12)), q, t, n, d)
    if gid == 13:
        sw = np.zeros((d * d, d * d), dtype=complex)
        for a in range(d):
            for b in range(d):
                sw[d * b + a, d * a + b] = 1
        return apply_2q(psi, sw, q, t, n, d)
    raise KeyError(gid)


def outcome_probs(psi, q, n, d):
    p = np.abs(psi.reshape((d,) * n)) ** 2
    axes = tuple(k for k in range(n) if k != q)
    return p.sum(axis=axes)


def project(psi, q, m, n, d):
    psi = psi.reshape((d,) * n).copy()
    idx = [slice(None)] * n
    for v in range(d):
        if v != m:
            idx[q] = v
            psi[tuple(idx)] = 0
    psi = psi.reshape(-1)
    return psi / np.linalg.norm(psi)


def pauli_matrix(xcol, zcol, phase, d):
    """tau^phase * prod_k X^x_k Z^z_k as a dense d^n x d^n matrix."""
    po = 2 if d == 2 else 1
    tau = np.exp(2j * np.pi / (po * d))
    X = one_qudit("X", d)
    Z = one_qudit("Z", d)
    m = np.array([[1.0 + 0j]])
    for xk, zk in zip(xcol, zcol):
        f = np.linalg.matrix_power(X, int(xk) % d) @ np.linalg.matrix_power(Z, int(zk) % d)
        m = np.kron(m, f)
    return (tau ** (int(phase) % (po * d))) * m


def apply_pauli_state(psi, xcol, zcol, d):
    n = len(xcol)
    X = one_qudit("X", d)
    Z = one_qudit("Z", d)
    for k in range(n):
        f = np.linalg.matrix_power(X, int(xcol[k]) % d) @ np.linalg.matrix_power(Z, int(zcol[k]) % d)
        psi = apply_1q(psi, f, k, n, d)
    return psi


def stabilizes(psi, xcol, zcol, phase, d, atol=1e-9):
    """True when tau^phase X^x Z^z |psi> == |psi>."""
    po = 2 if d == 2 else 1
    tau = np.exp(2j * np.pi / (po * d))
    out = (tau ** (int(phase) % (po * d))) * apply_pauli_state(psi, xcol, zcol, d)
    return np.allclose(out, psi, atol=atol)


def symplectic(x1, z1, x2, z2, d):
    """Commutation exponent k with A B = w^k B A for A = X^x1 Z^z1, B = X^x2 Z^z2."""
    return int((np.dot(z1, x2) - np.dot(x1, z2)) % d)


def outcome_distribution(n, d, ops, tol=1e-9):
    """Exact joint distribution of the measurement records of a noiseless circuit:
    {tuple of values per measurement event: probability}, by branching on every
    outcome of every M, M_X (H^-1 then a Z measurement, tableau_gates.py:280-296)
    and RESET (measure, then X^(-m), program.py:334-339)."""
    psi0 = np.zeros(d ** n, dtype=complex)
    psi0[0] = 1.0
    out = {}

    def walk(psi, k, prob, rec):
        while k < len(ops):
            gid, q, t = (int(v) for v in ops[k])
            if gid in (14, 15, 16):
                if gid == 15:
                    psi = apply_gate(psi, 6, q, -1, n, d)
                p = outcome_probs(psi, q, n, d)
                for m in range(d):
                    if p[m] > tol:
                        nxt = project(psi, q, m, n, d)
                        if gid == 16:
                            for _ in range((-m) % d):
                                nxt = apply_gate(nxt, 1, q, -1, n, d)
                        walk(nxt, k + 1, prob * p[m], rec + (m,))
                return
            if gid not in (0, 17):
                psi = apply_gate(psi, gid, q, t, n, d)
            k += 1
        out[rec] = out.get(rec, 0.0) + prob

    walk(psi0, 0, 1.0, ())
    return out
