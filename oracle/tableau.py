"""Literal CPU restatement of sdim's prime-dimension extended tableau.

Follows, step for step:
  * storage / initial state   sdim/tableau/dataclasses.py:30-39, tableau_prime.py:76-86
  * order / phase_order       sdim/tableau/dataclasses.py:78-106
  * modulo                    sdim/tableau/dataclasses.py:41-47, tableau_prime.py:88-95
  * gate sweeps               sdim/tableau/tableau_optimized.py:5-117
  * measure / pivot / exp.    sdim/tableau/tableau_prime.py:262-292, 365-380
  * random measurement        sdim/tableau/tableau_prime.py:294-334
  * deterministic measurement sdim/tableau/tableau_prime.py:336-363

Conventions kept from the reference: generators are COLUMNS, ``x[qudit, gen]``;
int64 numpy arrays; values are reduced lazily (every 64 ops by the driver and
inside ``measure``), Python floor semantics for ``%`` and ``//``.  The random
outcome of ``_random_measurement`` is supplied by the caller (a Philox draw)
instead of ``random.choice``.

The per-gate loops are interpreted Python loops over the generator index, the
same cost structure as the shipped reference with its numba JIT disabled
(tableau_optimized.py:4).  This file is test infrastructure only.
"""
import numpy as np


class OracleTableau:
    def __init__(self, n, d):
        self.n = int(n)
        self.d = int(d)
        self.even = (self.d % 2 == 0)
        self.po = 2 if self.even else 1            # phase_order, dataclasses.py:98-106
        self.order = 2 * self.d if self.even else self.d   # order, dataclasses.py:88-96
        n = self.n
        # stabilizers Z_i, destabilizers X_i, zero phases (dataclasses.py:30-39, tableau_prime.py:76-86)
        self.ph = np.zeros(n, dtype=np.int64)
        self.z = np.eye(n, dtype=np.int64)
        self.x = np.zeros((n, n), dtype=np.int64)
        self.dph = np.zeros(n, dtype=np.int64)
        self.dz = np.zeros((n, n), dtype=np.int64)
        self.dx = np.eye(n, dtype=np.int64)

    def copy(self):
        t = OracleTableau.__new__(OracleTableau)
        t.__dict__.update(self.__dict__)
        for k in ("ph", "z", "x", "dph", "dz", "dx"):
            setattr(t, k, getattr(self, k).copy())
        return t

    # ------------------------------------------------------------------ reduction
    def modulo(self):
        """dataclasses.py:41-47 + tableau_prime.py:88-95."""
        self.z %= self.d
        self.x %= self.d
        self.ph %= self.order
        self.dx %= self.d
        self.dz %= self.d
        self.dph %= self.order

    # ------------------------------------------------------------------ gate sweeps
    def hadamard(self, q):
        """tableau_optimized.py:5-30: (x, z) <- (-z, x); phase += po * x' * z'."""
        po = self.po
        for xs, zs, ps in ((self.x[q], self.z[q], self.ph), (self.dx[q], self.dz[q], self.dph)):
            for g in range(self.n):
                nx = -zs[g]
                nz = xs[g]
                ps[g] += po * nx * nz
                xs[g] = nx
                zs[g] = nz

    def hadamard_inv(self, q):
        """tableau_optimized.py:33-58: (x, z) <- (z, -x); phase += po * x' * z'."""
        po = self.po
        for xs, zs, ps in ((self.x[q], self.z[q], self.ph), (self.dx[q], self.dz[q], self.dph)):
            for g in range(self.n):
                nx = zs[g]
                nz = -xs[g]
                ps[g] += po * nx * nz
                xs[g] = nx
                zs[g] = nz

    def phase(self, q):
        """tableau_optimized.py:62-77: even: phase += x^2; odd: phase += x(x-1)/2; z += x."""
        for xs, zs, ps in ((self.x[q], self.z[q], self.ph), (self.dx[q], self.dz[q], self.dph)):
            for g in range(self.n):
                v = xs[g]
                if self.even:
                    ps[g] += v * v
                else:
                    ps[g] += (v * (v - 1)) // 2
                zs[g] += v

    def phase_inv(self, q):
        """tableau_optimized.py:81-96: the same increments subtracted; z -= x."""
        for xs, zs, ps in ((self.x[q], self.z[q], self.ph), (self.dx[q], self.dz[q], self.dph)):
            for g in range(self.n):
                v = xs[g]
                if self.even:
                    ps[g] -= v * v
                else:
                    ps[g] -= (v * (v - 1)) // 2
                zs[g] -= v

    def cnot(self, c, t):
        """tableau_optimized.py:99-107: x[t] += x[c]; z[c] += (d-1) z[t]; no phase."""
        dm1 = self.d - 1
        for xb, zb in ((self.x, self.z), (self.dx, self.dz)):
            for g in range(self.n):
                xb[t, g] += xb[c, g]
                zb[c, g] += zb[t, g] * dm1

    def cnot_inv(self, c, t):
        """tableau_optimized.py:110-117: x[t] -= x[c]; z[c] -= (d-1) z[t]."""
        dm1 = self.d - 1
        for xb, zb in ((self.x, self.z), (self.dx, self.dz)):
            for g in range(self.n):
                xb[t, g] -= xb[c, g]
                zb[c, g] -= zb[t, g] * dm1

    # ------------------------------------------------------------------ measurement
    def exponentiate(self, col, e):
        """tableau_prime.py:365-380: (X^a Z^b)^e = w^(ab e(e-1)/2) X^(ea) Z^(eb)."""
        self.ph[col] *= e
        self.ph[col] += int(np.dot(self.x[:, col], self.z[:, col])) * e * (e - 1) // 2 * self.po
        self.x[:, col] *= e
        self.z[:, col] *= e
        self.ph[col] %= self.order

    def measure(self, q, outcome_source):
        """tableau_prime.py:262-292.  ``outcome_source()`` is called only for a random outcome.

        Returns (deterministic: bool, value: int).
        """
        d = self.d
        pivot = None
        for i in range(self.n):                       # first stabilizer with x[q, i] != 0 (:275-283)
            xp = int(self.x[q, i]) % d
            if xp > 0:
                pivot = i
                if xp != 1:
                    self.exponentiate(pivot, pow(xp, -1, d))
                break
        self.x %= d                                   # :284-289
        self.z %= d
        self.ph %= self.order
        self.dx %= d
        self.dz %= d
        self.dph %= self.order
        if pivot is not None:
            return self._random(q, pivot, outcome_source)
        return self._deterministic(q)

    def _random(self, q, p, outcome_source):
        """tableau_prime.py:294-334."""
        d, po, order = self.d, self.po, self.order
        xp = self.x[:, p]
        zp = self.z[:, p]
        xz = int(np.dot(xp, zp))
        for i in range(self.n):
            if self.dx[q, i] != 0:                    # destabilizer column i (:309-315)
                f = int(-self.dx[q, i]) % d
                cp = int(np.dot(self.dz[:, i], xp * f))
                cp += xz * f * (f - 1) // 2 * po
                self.dx[:, i] = (self.dx[:, i] + xp * f) % d
                self.dz[:, i] = (self.dz[:, i] + zp * f) % d
                self.dph[i] = (self.dph[i] + self.ph[p] * f + po * cp) % order
            if self.x[q, i] != 0 and i != p:           # stabilizer column i (:316-322)
                f = int(-self.x[q, i]) % d
                cp = int(np.dot(self.z[:, i], xp * f))
                cp += xz * f * (f - 1) // 2 * po
                self.x[:, i] = (self.x[:, i] + xp * f) % d
                self.z[:, i] = (self.z[:, i] + zp * f) % d
                self.ph[i] = (self.ph[i] + self.ph[p] * f + po * cp) % order
        self.dx[:, p] = self.x[:, p]                  # destabilizer p <- stabilizer p (:325-327)
        self.dz[:, p] = self.z[:, p]
        self.dph[p] = self.ph[p]
        self.z[:, p] = 0                              # stabilizer p <- Z_q (:329-331)
        self.z[q, p] = 1
        self.x[:, p] = 0
        m = int(outcome_source())
        self.ph[p] = (-m * po) % order                # :332-333
        return False, m

    def _deterministic(self, q):
        """tableau_prime.py:336-363: accumulate prod_i S_i^{f_i}, f_i = dx[q, i]."""
        d, po, order = self.d, self.po, self.order
        n = self.n
        ax = np.zeros(n, dtype=np.int64)
        az = np.zeros(n, dtype=np.int64)
        ap = 0
        for i in range(n):
            f = int(self.dx[q, i]) % d
            if f != 0:
                cp = int(np.dot(az, f * self.x[:, i]))
                cp += int(np.dot(self.x[:, i], self.z[:, i])) * f * (f - 1) // 2 * po
                ax += self.x[:, i] * f
                az += self.z[:, i] * f
                ap += f * int(self.ph[i]) + po * cp
        ap %= order
        return True, (-ap // po) % d

    # ------------------------------------------------------------------ export
    def arrays(self):
        """(x, z, ph, dx, dz, dph) copies in the reference layout."""
        return (self.x.copy(), self.z.copy(), self.ph.copy(),
                self.dx.copy(), self.dz.copy(), self.dph.copy())
