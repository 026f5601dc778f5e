"""Pins the CPU oracle (oracle/) before anything is compared against it.

Every check here is independent of the oracle's own formulas:
  * Philox4x64-10 against numpy.random.Philox (a known-answer generator);
  * the deterministic outcomes the reference's own tests assert
    (tests/test_circuit.py:9-263 prime cases, tests/test_program.py:14-47,
    tests/test_noise_and_io.py:174-200, tests/test_reset.py:11-36);
  * a dense statevector built from sdim/unitary.py's matrix formulas: every
    stabilizer (with its exact phase) fixes the state after every op,
    deterministic outcomes are certain, random outcomes have probability 1/d;
  * operator algebra on dense matrices: each gate conjugates every stabilizer
    AND destabilizer column (phase included); each random-measurement rowsum
    equals the matrix product D_i S_p^f; the deterministic ancilla equals
    tau^ap Z_q; the stabilizer/destabilizer symplectic pairing is kept;
  * the X/Z closed forms against the reference's literal H/P words;
  * the literal Python oracle against the C restatement (lazy vs eager).
"""
import numpy as np
import pytest

from oracle import c_oracle, philox
from oracle import gates as OG
from oracle import program as OP
from oracle import statevector as SV
from oracle.tableau import OracleTableau


# ----------------------------------------------------------------------------- Philox
@pytest.mark.parametrize("key,ctr", [((0, 0), (0, 0, 0, 0)), ((123, 456), (7, 8, 9, 10)),
                                     ((2**64 - 1, 5), (2**64 - 2, 3, 2**63, 1))])
def test_philox_matches_numpy(key, ctr):
    bg = np.random.Philox(key=np.array(key, dtype=np.uint64), counter=np.array(ctr, dtype=np.uint64))
    raw = [int(bg.random_raw()) for _ in range(8)]
    v = sum(int(x) << (64 * i) for i, x in enumerate(ctr))
    blocks = []
    for step in (1, 2):  # numpy increments the 256-bit counter before each block
        w = (v + step) % (1 << 256)
        c = [(w >> (64 * i)) & philox.MASK64 for i in range(4)]
        blocks += list(philox.philox4x64_10(c, key))
    assert raw == blocks


def test_draw_ranges():
    for d in (2, 3, 5, 7, 11):
        seen = {philox.measurement_outcome(1, 2, op, 0, d) for op in range(400)}
        assert seen == set(range(d))
    assert philox.keep_threshold53(0.0) == 1 << 53
    assert philox.keep_threshold53(1.0) == 0
    # p = 1 always applies, p = 0 never applies
    for op in range(50):
        assert philox.noise_exponents(3, 4, op, 0, 5, 0, 0) != (0, 0)
        assert philox.noise_exponents(3, 4, op, 0, 5, 1, 1 << 53) == (0, 0)


# ----------------------------------------------------------------------------- reference expectations
def _run(n, d, ops, shots=1, **kw):
    recs, tabs = OP.simulate(n, d, ops, shots=shots, **kw)
    return [(q, det, v) for (q, r, s, det, v) in recs], tabs


def test_reference_phase_kickback():      # tests/test_circuit.py:9-27
    ops = [(5, 1, -1), (7, 1, -1), (5, 0, -1), (9, 0, 1), (5, 1, -1), (9, 0, 1), (5, 1, -1),
           (7, 0, -1), (5, 0, -1), (14, 0, -1), (7, 1, -1), (5, 1, -1), (14, 1, -1)]
    assert _run(2, 2, ops)[0] == [(0, True, 1), (1, True, 1)]


def test_reference_qubit_and_qutrit_flip():   # tests/test_circuit.py:29-46
    ops = [(5, 0, -1), (7, 0, -1), (7, 0, -1), (5, 0, -1), (14, 0, -1), (1, 1, -1), (14, 1, -1)]
    assert _run(2, 2, ops)[0] == [(0, True, 1), (1, True, 1)]
    assert _run(2, 3, [(1, 0, -1), (14, 0, -1)])[0] == [(0, True, 1)]


@pytest.mark.parametrize("d", [2, 3, 5, 7])
def test_reference_swap_computational_basis(d):   # tests/test_circuit.py:48-94 (prime d)
    for x0 in range(d):
        for x1 in range(d):
            ops = [(1, 0, -1)] * x0 + [(1, 1, -1)] * x1 + [(13, 0, 1), (14, 0, -1), (14, 1, -1)]
            assert _run(2, d, ops)[0] == [(0, True, x1), (1, True, x0)]
            ops = [(1, 0, -1)] * x0 + [(1, 1, -1)] * x1 + [(13, 0, 1), (13, 0, 1), (14, 0, -1), (14, 1, -1)]
            assert _run(2, d, ops)[0] == [(0, True, x0), (1, True, x1)]


@pytest.mark.parametrize("a", range(3))
@pytest.mark.parametrize("b", range(3))
def test_reference_qutrit_swap_x_basis(a, b):    # tests/test_circuit.py:96-136
    ops = [(1, 0, -1)] * a + [(5, 0, -1)] + [(1, 1, -1)] * b + [(5, 1, -1), (13, 0, 1), (15, 0, -1), (15, 1, -1)]
    assert _run(2, 3, ops)[0] == [(0, True, b), (1, True, a)]


@pytest.mark.parametrize("f0,f1", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_reference_qubit_deutsch(f0, f1):        # tests/test_circuit.py:138-167
    ops = [(5, 0, -1), (1, 1, -1), (5, 1, -1)]
    if f0:
        ops += [(9, 0, 1), (1, 1, -1)]
    if f1:
        ops += [(9, 0, 1)]
    ops += [(5, 0, -1), (14, 0, -1)]
    assert _run(2, 2, ops)[0] == [(0, True, 0 if f0 == f1 else 1)]


def test_reference_stabilizer_extraction():      # tests/test_circuit.py:169-222
    ops = [(1, 2, -1), (1, 3, -1), (1, 4, -1), (5, 0, -1), (11, 0, 2), (11, 0, 3), (12, 0, 4), (6, 0, -1),
           (5, 1, -1), (12, 1, 2), (12, 1, 3), (11, 1, 4), (6, 1, -1), (14, 0, -1), (14, 1, -1)]
    assert _run(5, 3, ops)[0] == [(0, True, 1), (1, True, 2)]
    ops = [(5, 2, -1), (5, 3, -1), (5, 4, -1), (3, 2, -1), (5, 0, -1), (9, 0, 2), (9, 0, 3), (10, 0, 4),
           (6, 0, -1), (5, 1, -1), (10, 1, 2), (10, 1, 3), (9, 1, 4), (6, 1, -1), (14, 0, -1), (14, 1, -1)]
    assert _run(5, 3, ops)[0] == [(0, True, 2), (1, True, 1)]


def test_reference_deutsch_qutrit():             # tests/test_circuit.py:224-263
    ops = [(5, 0, -1), (1, 1, -1), (5, 1, -1), (9, 0, 1), (6, 0, -1), (14, 0, -1)]
    assert _run(2, 3, ops)[0] == [(0, True, 2)]


def test_reference_measurement_format_values():  # tests/test_program.py:14-47
    recs, _ = OP.simulate(2, 3, [(14, 0, -1), (14, 1, -1), (14, 0, -1), (14, 1, -1)], shots=2)
    assert [(q, r, s, det, v) for (q, r, s, det, v) in recs] == [
        (0, 0, 0, True, 0), (1, 0, 0, True, 0), (0, 1, 0, True, 0), (1, 1, 0, True, 0),
        (0, 0, 1, True, 0), (1, 0, 1, True, 0), (0, 1, 1, True, 0), (1, 1, 1, True, 0)]


@pytest.mark.parametrize("d", [3, 5])
def test_reference_identity_rb_all_zero(d):      # tests/test_noise_and_io.py:174-200
    rng = np.random.default_rng(d)
    inv = {0: 0, 1: 2, 2: 1, 3: 4, 4: 3, 5: 6, 6: 5, 7: 8, 8: 7}
    seq = [int(g) for g in rng.integers(0, 9, size=200)]
    ops = [(g, 0, -1) for g in seq] + [(inv[g], 0, -1) for g in reversed(seq)] + [(14, 0, -1)]
    recs, _ = c_oracle.simulate(1, d, ops, shots=64, seed=5)
    assert (recs[:, 0, 4] == 0).all()


@pytest.mark.parametrize("d", [3, 5, 7, 11])
def test_reference_reset_then_x(d):              # tests/test_reset.py:11-36
    p = (d * d - 1) / (d * d)
    for num_x in range(d):
        ops = [(17, 0, -1), (16, 0, -1)] + [(1, 0, -1)] * num_x + [(14, 0, -1)]
        recs, _ = c_oracle.simulate(1, d, ops, shots=200, seed=11, stream=num_x,
                                    noise={0: (0, philox.keep_threshold53(p))})
        assert (recs[:, 1, 4] == num_x).all()
        assert recs[:, 1, 3].all()


# ----------------------------------------------------------------------------- statevector
def _rand_circuit(rng, n, d, depth, meas=0.2, noise=0.1):
    ops, nz = [], {}
    for _ in range(depth):
        r = rng.random()
        if r < meas:
            ops.append((int(rng.choice([14, 15, 16])), int(rng.integers(n)), -1))
        elif r < meas + noise:
            ops.append((17, int(rng.integers(n)), -1))
            nz[len(ops) - 1] = (int(rng.integers(3)), philox.keep_threshold53(float(rng.random())))
        elif r < 0.65 or n < 2:
            ops.append((int(rng.integers(1, 9)), int(rng.integers(n)), -1))
        else:
            a, b = rng.choice(n, 2, replace=False)
            ops.append((int(rng.integers(9, 14)), int(a), int(b)))
    return ops, nz


def _check_stabilized(t, psi, d):
    for g in range(t.n):
        assert SV.stabilizes(psi, t.x[:, g], t.z[:, g], t.ph[g], d), f"stabilizer {g} broken"


@pytest.mark.parametrize("d", [2, 3, 5])
def test_oracle_against_statevector(d):
    rng = np.random.default_rng(100 + d)
    n = 3 if d < 5 else 2
    for trial in range(25):
        ops, nz = _rand_circuit(rng, n, d, 40)
        t = OracleTableau(n, d)
        psi = np.zeros(d ** n, dtype=complex)
        psi[0] = 1.0
        for k, (gid, q, tg) in enumerate(ops):
            if gid in OG.UNITARY_GATES:
                OG.UNITARY_GATES[gid](t, q, tg)
                psi = SV.apply_gate(psi, gid, q, tg, n, d)
            elif gid == 17:
                a, b = philox.noise_exponents(9, trial, k, 0, d, *nz[k])
                for _ in range(a):
                    OG.apply_X(t, q)
                for _ in range(b):
                    OG.apply_Z(t, q)
                psi = SV.apply_1q(psi, np.linalg.matrix_power(SV.one_qudit("X", d), a)
                                  @ np.linalg.matrix_power(SV.one_qudit("Z", d), b), q, n, d)
            else:
                if gid == 15:
                    t.hadamard_inv(q)
                    psi = SV.apply_gate(psi, 6, q, -1, n, d)
                probs = SV.outcome_probs(psi, q, n, d)
                det, m = t.measure(q, lambda: philox.measurement_outcome(9, trial, k, 0, d))
                if det:
                    assert probs[m] == pytest.approx(1.0, abs=1e-9)
                else:
                    assert np.allclose(probs, 1.0 / d, atol=1e-9)
                psi = SV.project(psi, q, m, n, d)
                if gid == 16:
                    for _ in range((-m) % d):
                        OG.apply_X(t, q)
                    psi = SV.apply_1q(psi, np.linalg.matrix_power(SV.one_qudit("X", d), (-m) % d), q, n, d)
                    assert SV.outcome_probs(psi, q, n, d)[0] == pytest.approx(1.0, abs=1e-9)
            _check_stabilized(t, psi, d)


# ----------------------------------------------------------------------------- operator algebra
def _cols(t):
    stab = [(t.x[:, g].copy(), t.z[:, g].copy(), int(t.ph[g])) for g in range(t.n)]
    dest = [(t.dx[:, g].copy(), t.dz[:, g].copy(), int(t.dph[g])) for g in range(t.n)]
    return stab, dest


def _mat(col, d):
    return SV.pauli_matrix(col[0], col[1], col[2], d)


def _unitary(gid, q, tg, n, d):
    dim = d ** n
    u = np.zeros((dim, dim), dtype=complex)
    for i in range(dim):
        e = np.zeros(dim, dtype=complex)
        e[i] = 1
        u[:, i] = SV.apply_gate(e, gid, q, tg, n, d)
    return u


def _random_tableau(rng, n, d, depth=30):
    t = OracleTableau(n, d)
    for k in range(depth):
        if rng.random() < 0.15:
            t.measure(int(rng.integers(n)), lambda: int(rng.integers(d)))
        elif rng.random() < 0.6 or n < 2:
            OG.UNITARY_GATES[int(rng.integers(1, 9))](t, int(rng.integers(n)), -1)
        else:
            a, b = rng.choice(n, 2, replace=False)
            OG.UNITARY_GATES[int(rng.integers(9, 14))](t, int(a), int(b))
    return t


@pytest.mark.parametrize("d", [2, 3, 5])
def test_gates_conjugate_every_column(d):
    """U C U^dag == C' for every stabilizer and destabilizer column, phases included."""
    rng = np.random.default_rng(7 + d)
    n = 2
    for trial in range(12):
        t = _random_tableau(rng, n, d)
        for gid in range(1, 14):
            q, tg = (0, 1) if rng.random() < 0.5 else (1, 0)
            before = _cols(t)
            t2 = t.copy()
            OG.UNITARY_GATES[gid](t2, q, tg)
            after = _cols(t2)
            u = _unitary(gid, q, tg, n, d)
            for blk in (0, 1):
                for c0, c1 in zip(before[blk], after[blk]):
                    assert np.allclose(u @ _mat(c0, d) @ u.conj().T, _mat(c1, d), atol=1e-9), (gid, blk)


def _symplectic_ok(t):
    d, n = t.d, t.n
    stab, dest = _cols(t)
    for i in range(n):
        for j in range(n):
            assert SV.symplectic(stab[i][0], stab[i][1], stab[j][0], stab[j][1], d) == 0
            assert SV.symplectic(dest[i][0], dest[i][1], dest[j][0], dest[j][1], d) == 0
            k = SV.symplectic(dest[i][0], dest[i][1], stab[j][0], stab[j][1], d)
            assert (k != 0) == (i == j)


@pytest.mark.parametrize("d", [2, 3, 5, 7])
def test_random_measurement_rowsum_is_operator_product(d):
    """tableau_prime.py:294-334: new col_i == old col_i * (stab_p)^f_i as matrices (exact phase)."""
    rng = np.random.default_rng(31 + d)
    n = 2 if d > 3 else 3
    done = 0
    for trial in range(200):
        t = _random_tableau(rng, n, d, depth=25)
        t.modulo()
        _symplectic_ok(t)
        q = int(rng.integers(n))
        piv = [i for i in range(n) if t.x[q, i] % d]
        if not piv:
            continue
        p = piv[0]
        t_exp = t.copy()
        xv = int(t_exp.x[q, p]) % d
        if xv != 1:
            t_exp.exponentiate(p, pow(xv, -1, d))
            # exponentiation is the operator power S_p^e
            assert np.allclose(np.linalg.matrix_power(_mat(_cols(t)[0][p], d), pow(xv, -1, d)),
                               _mat(_cols(t_exp)[0][p], d), atol=1e-9)
        t_exp.modulo()
        sp = _mat(_cols(t_exp)[0][p], d)
        stab0, dest0 = _cols(t_exp)
        t2 = t.copy()
        m = int(rng.integers(d))
        det, mm = t2.measure(q, lambda: m)
        assert not det and mm == m
        stab1, dest1 = _cols(t2)
        for i in range(n):
            f = (-int(stab0[i][0][q])) % d
            if i != p:
                assert np.allclose(_mat(stab0[i], d) @ np.linalg.matrix_power(sp, f), _mat(stab1[i], d), atol=1e-9)
                f = (-int(dest0[i][0][q])) % d
                assert np.allclose(_mat(dest0[i], d) @ np.linalg.matrix_power(sp, f), _mat(dest1[i], d), atol=1e-9)
        assert np.allclose(_mat(dest1[p], d), sp, atol=1e-9)          # destab p <- stab p
        zq = np.zeros(n, dtype=np.int64)
        zq[q] = 1
        po = 2 if d == 2 else 1
        assert np.allclose(_mat(stab1[p], d), _mat((np.zeros(n, dtype=np.int64), zq, -m * po), d), atol=1e-9)
        _symplectic_ok(t2)
        done += 1
    assert done > 20


@pytest.mark.parametrize("d", [2, 3, 5, 7])
def test_deterministic_ancilla_is_tau_ap_Zq(d):
    """tableau_prime.py:336-363: prod_i S_i^{f_i} == tau^ap Z_q and m == -ap/po mod d."""
    rng = np.random.default_rng(55 + d)
    n = 2 if d > 3 else 3
    done = 0
    po = 2 if d == 2 else 1
    tau = np.exp(2j * np.pi / (po * d))
    for trial in range(150):
        t = _random_tableau(rng, n, d, depth=25)
        t.modulo()
        q = int(rng.integers(n))
        if any(t.x[q, i] % d for i in range(n)):
            continue
        stab, dest = _cols(t)
        prod = np.eye(d ** n, dtype=complex)
        for i in range(n):
            f = int(t.dx[q, i]) % d
            prod = prod @ np.linalg.matrix_power(_mat(stab[i], d), f)
        zq = np.zeros(n, dtype=np.int64)
        zq[q] = 1
        zmat = _mat((np.zeros(n, dtype=np.int64), zq, 0), d)
        det, m = t.copy().measure(q, lambda: 0)
        assert det
        # the product is tau^ap Z_q with ap = -m*po (mod po*d)
        assert np.allclose(prod, (tau ** ((-m * po) % (po * d))) * zmat, atol=1e-9)
        done += 1
    assert done > 10


@pytest.mark.parametrize("d", [2, 3, 5, 7, 11])
def test_pauli_closed_forms_match_literal_words(d):
    """X: ph -= po z, X^-1: ph += po z, Z: ph += po x, Z^-1: ph -= po x; blocks unchanged."""
    rng = np.random.default_rng(d)
    n = 4
    po = 2 if d == 2 else 1
    for trial in range(10):
        t = _random_tableau(rng, n, d, depth=20)
        t.modulo()
        q = int(rng.integers(n))
        for fn, sign, blk in ((OG.apply_X, -1, "z"), (OG.apply_X_inv, +1, "z"),
                              (OG.apply_Z, +1, "x"), (OG.apply_Z_inv, -1, "x")):
            t2 = t.copy()
            fn(t2, q)
            t2.modulo()
            assert np.array_equal(t2.x, t.x) and np.array_equal(t2.z, t.z)
            assert np.array_equal(t2.dx, t.dx) and np.array_equal(t2.dz, t.dz)
            row = getattr(t, blk)[q]
            drow = getattr(t, "d" + blk)[q]
            assert np.array_equal(t2.ph, (t.ph + sign * po * row) % t.order)
            assert np.array_equal(t2.dph, (t.dph + sign * po * drow) % t.order)


@pytest.mark.parametrize("d", [2, 3, 5, 7, 11])
def test_cz_and_swap_closed_forms_match_literal_words(d):
    """The engine's fused two-row ops: CZ^(+-1)(c, t): z[t] +-= x[c], z[c] +-= x[t],
    ph +-= po x[c] x[t]; SWAP: rows exchanged, phases unchanged (tableau_gates.py:229-261, 319-328)."""
    rng = np.random.default_rng(40 + d)
    n = 5
    po = 2 if d == 2 else 1
    for trial in range(10):
        t = _random_tableau(rng, n, d, depth=30)
        t.modulo()
        c, tg = (int(v) for v in rng.choice(n, 2, replace=False))
        for blocks in (("x", "z", "ph"), ("dx", "dz", "dph")):
            X, Z, P = (getattr(t, b) for b in blocks)
            for fn, sign in ((OG.apply_CZ, 1), (OG.apply_CZ_inv, -1)):
                t2 = t.copy()
                fn(t2, c, tg)
                t2.modulo()
                ez = Z.copy()
                ez[tg] = (Z[tg] + sign * X[c]) % d
                ez[c] = (Z[c] + sign * X[tg]) % d
                assert np.array_equal(getattr(t2, blocks[0]), X)
                assert np.array_equal(getattr(t2, blocks[1]), ez)
                assert np.array_equal(getattr(t2, blocks[2]), (P + sign * po * X[c] * X[tg]) % t.order)
            t2 = t.copy()
            OG.apply_SWAP(t2, c, tg)
            t2.modulo()
            for b, A in zip(blocks[:2], (X, Z)):
                e = A.copy()
                e[[c, tg]] = e[[tg, c]]
                assert np.array_equal(getattr(t2, b), e)
            assert np.array_equal(getattr(t2, blocks[2]), P)


# ----------------------------------------------------------------------------- Python vs C
@pytest.mark.parametrize("d", [2, 3, 5, 7, 11])
def test_python_oracle_matches_c_restatement(d):
    rng = np.random.default_rng(1000 + d)
    for trial in range(25):
        n = int(rng.integers(1, 7))
        ops, nz = _rand_circuit(rng, n, d, int(rng.integers(1, 70)), meas=0.15)
        r1, t1 = OP.simulate(n, d, ops, shots=2, seed=3, stream=trial, noise=nz)
        r2, t2 = c_oracle.simulate(n, d, ops, shots=2, seed=3, stream=trial, noise=nz, keep_shots=2)
        r1 = np.array(r1, dtype=np.int64).reshape(2, -1, 5)
        assert np.array_equal(r1, r2)
        for a, b in zip(t1, t2):
            for u, v in zip(a, b):
                assert np.array_equal(u, v)
