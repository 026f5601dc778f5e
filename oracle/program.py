"""Shot loop of the oracle, restating ``Program._simulate_tableau``.

Follows sdim/program.py:269-365 (tableau path, ``force_tableau`` semantics):
per shot a fresh initial tableau (:309), ``modulo`` every 64 ops (:317-318),
``apply_gate`` dispatch (:367-385), per-qudit round counters (:321-332), the
RESET fix-up that applies X ``(-m) mod d`` times (:334-339), and a final
``modulo`` (:351).

Two deliberate differences, shared with the device engine:
  * the random outcome is a Philox draw keyed by (seed, stream, op_index, shot)
    instead of ``random.choice`` (tableau_prime.py:332);
  * N1 (gate 17) is applied to the tableau as X^a Z^b using the reference's
    X / Z words (tableau_gates.py:27-110), with the channel semantics of
    ``_build_ir`` (program.py:484-510); the reference's tableau path maps N1 to
    the identity (program.py:31) and only its frame sampler draws noise.

Input op list: tuples (gate_id, qudit, target) with target -1 for one-qudit
gates; ``noise`` maps op_index -> (channel_code, keep53) for every N1 op.
"""
from . import gates as G
from . import philox
from .tableau import OracleTableau


def simulate(n, d, ops, shots=1, seed=0, stream=0, noise=None, keep_tableaux=True,
             initial=None):
    """Run the circuit ``shots`` times.

    Returns (records, tableaux):
      records  list of (qudit, round, shot, deterministic, value) in shot-major, time order
      tableaux list (one per shot) of (x, z, ph, dx, dz, dph) int64 arrays, reduced
    """
    noise = noise or {}
    records = []
    tableaux = []
    for shot in range(shots):
        t = initial.copy() if initial is not None else OracleTableau(n, d)
        counts = [0] * n
        for time, op in enumerate(ops):
            gid, q, tg = int(op[0]), int(op[1]), int(op[2])
            if time % 64 == 0:
                t.modulo()
            if gid == 0:
                continue
            if gid in G.UNITARY_GATES:
                G.UNITARY_GATES[gid](t, q, tg)
                continue
            if gid == G.NOISE:
                channel, keep53 = noise[time]
                a, b = philox.noise_exponents(seed, stream, time, shot, d, channel, keep53)
                for _ in range(a):
                    G.apply_X(t, q)
                for _ in range(b):
                    G.apply_Z(t, q)
                continue
            if gid in (G.MEASURE, G.MEASURE_X, G.RESET):
                if gid == G.MEASURE_X:                # tableau_gates.py:280-296: H^-1 then M
                    t.hadamard_inv(q)
                det, m = t.measure(q, lambda: philox.measurement_outcome(seed, stream, time, shot, d))
                records.append((q, counts[q], shot, bool(det), int(m)))
                counts[q] += 1
                if gid == G.RESET:                    # program.py:334-339
                    for _ in range((-m) % d):
                        G.apply_X(t, q)
                continue
            raise ValueError("Invalid gate value")
        t.modulo()
        if keep_tableaux:
            tableaux.append(t.arrays())
    return records, tableaux
