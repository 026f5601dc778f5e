"""Host-side tableau types with the reference's fields and conventions.

``Tableau`` / ``ExtendedTableau`` hold the reference layout (generators are
columns: ``x_block[qudit, generator]``, int64; sdim/tableau/dataclasses.py:7-164,
sdim/tableau/tableau_prime.py:8-95).  They are what ``Program`` returns and
accepts as an initial state.  The gate and measurement methods of
``ExtendedTableau`` run on the GPU engine (one small job per call; the fast
path is ``Program.simulate``), there is no CPU simulation in this package.

``MeasurementResult`` is the reference's record type (dataclasses.py:166-210).
"""
import random
from dataclasses import dataclass
from functools import cached_property
from math import gcd
from typing import Optional

import numpy as np


@dataclass
class Tableau:
    num_qudits: int = 1
    dimension: int = 2
    phase_vector: Optional[np.ndarray] = None
    z_block: Optional[np.ndarray] = None
    x_block: Optional[np.ndarray] = None

    def __post_init__(self):
        n = self.num_qudits
        if self.phase_vector is None:
            self.phase_vector = np.zeros(n, dtype=np.int64)
        if self.z_block is None:
            self.z_block = np.eye(n, dtype=np.int64)
        if self.x_block is None:
            self.x_block = np.zeros((n, n), dtype=np.int64)

    def modulo(self):
        self.z_block %= self.dimension
        self.x_block %= self.dimension
        self.phase_vector %= self.order

    @cached_property
    def coprime_order(self) -> set:
        return {i for i in range(1, self.order) if gcd(i, self.order) == 1}

    @cached_property
    def coprime_dimension(self) -> set:
        return {i for i in range(1, self.dimension) if gcd(i, self.dimension) == 1}

    @cached_property
    def prime(self) -> bool:
        return is_prime(self.dimension)

    @property
    def even(self) -> bool:
        return self.dimension % 2 == 0

    @property
    def order(self) -> int:
        """Order of the phase group: 2d for even d, d for odd d (phases count w^(1/po))."""
        return 2 * self.dimension if self.even else self.dimension

    @property
    def phase_order(self) -> int:
        return 2 if self.even else 1

    @property
    def pauli_size(self) -> int:
        return 2 * self.num_qudits + 1

    @property
    def stab_tableau(self) -> np.ndarray:
        return np.vstack((self.phase_vector, self.z_block, self.x_block))

    def _print_labeled_matrix(self, label, matrix):
        print(f"{label}:")
        print(matrix)

    def print_phase_vector(self):
        self._print_labeled_matrix("Phase Vector", self.phase_vector)

    def print_z_block(self):
        self._print_labeled_matrix("Z Block", self.z_block)

    def print_x_block(self):
        self._print_labeled_matrix("X Block", self.x_block)

    def print_tableau(self):
        self.print_phase_vector()
        self.print_z_block()
        self.print_x_block()


def is_prime(d: int) -> bool:
    if d < 2:
        return False
    k = 2
    while k * k <= d:
        if d % k == 0:
            return False
        k += 1
    return True


@dataclass
class ExtendedTableau(Tableau):
    destab_phase_vector: Optional[np.ndarray] = None
    destab_z_block: Optional[np.ndarray] = None
    destab_x_block: Optional[np.ndarray] = None

    def __post_init__(self):
        super().__post_init__()
        n = self.num_qudits
        if self.destab_z_block is None:
            self.destab_z_block = np.zeros((n, n), dtype=np.int64)
        if self.destab_x_block is None:
            self.destab_x_block = np.eye(n, dtype=np.int64)
        if self.destab_phase_vector is None:
            self.destab_phase_vector = np.zeros(n, dtype=np.int64)

    def print_destab_phase_vector(self):
        self._print_labeled_matrix("Destabilizer Phase Vector", self.destab_phase_vector)

    def print_destab_z_block(self):
        self._print_labeled_matrix("Destabilizer Z Block", self.destab_z_block)

    def print_destab_x_block(self):
        self._print_labeled_matrix("Destabilizer X Block", self.destab_x_block)

    def print_tableau(self):
        super().print_tableau()
        self.print_destab_phase_vector()
        self.print_destab_z_block()
        self.print_destab_x_block()

    @property
    def destab_tableau(self) -> np.ndarray:
        return np.vstack((self.destab_phase_vector, self.destab_z_block, self.destab_x_block))

    @property
    def tableau(self) -> np.ndarray:
        return np.hstack((self.stab_tableau, self.destab_tableau))

    def modulo(self):
        super().modulo()
        self.destab_x_block %= self.dimension
        self.destab_z_block %= self.dimension
        self.destab_phase_vector %= self.order

    # ------------------------------------------------------------ engine-backed operations
    def arrays(self):
        """(x, z, ph, dx, dz, dph) in the order of the engine's initial-tableau argument."""
        return (self.x_block, self.z_block, self.phase_vector,
                self.destab_x_block, self.destab_z_block, self.destab_phase_vector)

    def set_arrays(self, arrs):
        (self.x_block, self.z_block, self.phase_vector,
         self.destab_x_block, self.destab_z_block, self.destab_phase_vector) = [np.array(a, dtype=np.int64) for a in arrs]

    def is_initial_state(self) -> bool:
        n = self.num_qudits
        d = self.dimension
        eye = np.eye(n, dtype=np.int64)
        return (np.array_equal(self.x_block % d, 0 * eye) and np.array_equal(self.z_block % d, eye)
                and np.array_equal(self.destab_x_block % d, eye) and np.array_equal(self.destab_z_block % d, 0 * eye)
                and not np.any(self.phase_vector % self.order) and not np.any(self.destab_phase_vector % self.order))

    def _run(self, ops, seed=None):
        from .engine import Job, default_engine
        self.modulo()
        ir = np.asarray(ops, dtype=np.int64).reshape(-1, 3)
        job = Job(self.num_qudits, self.dimension, ir, shots=1,
                  seed=random.getrandbits(63) if seed is None else seed, init=self.arrays())
        tk = default_engine().run([job])
        rec = tk.records(0)
        self.set_arrays(tk.tableau(0, 0))
        tk.release()
        return rec

    def hadamard(self, qudit_index: int):
        self._run([(5, qudit_index, -1)])

    def hadamard_inv(self, qudit_index: int):
        self._run([(6, qudit_index, -1)])

    def phase(self, qudit_index: int):
        self._run([(7, qudit_index, -1)])

    def phase_inv(self, qudit_index: int):
        self._run([(8, qudit_index, -1)])

    def cnot(self, control: int, target: int):
        self._run([(9, control, target)])

    def cnot_inv(self, control: int, target: int):
        self._run([(10, control, target)])

    def measure(self, qudit_index: int) -> "MeasurementResult":
        rec = self._run([(14, qudit_index, -1)])
        r = rec[0, 0]
        return MeasurementResult(int(r['qudit_index']), bool(r['deterministic']), int(r['measurement_value']))

    def exponentiate(self, col: int, exponent: int):
        """(X^a Z^b)^e = w^(a.b e(e-1)/2) X^(ea) Z^(eb) on stabilizer column ``col``
        (sdim/tableau/tableau_prime.py:365-380); a pure column operation on the host arrays."""
        e = int(exponent)
        xz = int(np.dot(self.x_block[:, col], self.z_block[:, col]))
        self.phase_vector[col] = (self.phase_vector[col] * e + xz * e * (e - 1) // 2 * self.phase_order) % self.order
        self.x_block[:, col] *= e
        self.z_block[:, col] *= e


@dataclass
class MeasurementResult:
    qudit_index: int
    deterministic: bool
    measurement_value: int
    stabilizer_tableau: Optional[Tableau] = None

    def __str__(self) -> str:
        kind = "deterministic" if self.deterministic else "random"
        return f"Measured qudit ({self.qudit_index}) as ({self.measurement_value}) and was {kind}"

    def __repr__(self) -> str:
        return str(self)

    def get_tableau(self):
        if self.stabilizer_tableau is None:
            raise ValueError("Stabilizer tableau not recorded during measurement")
        return self.stabilizer_tableau
