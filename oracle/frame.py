"""Oracle (test infrastructure only): restatement of sdim's Pauli-frame sampler.

Follows `simulate_frame` (sdim/program.py:45-165) and the frame branch of
`Program.simulate` (program.py:244-262): shot 0 is a noiseless tableau shot
(the reference's tableau path treats N1 as the identity, program.py:31) whose
records are the reference results; every further shot propagates a Pauli frame:

* frames start with x = 0 and z uniform in [0, d) on every qudit (program.py:63-64);
* gates conjugate the frame (program.py:79-123): X/Z and their inverses do
  nothing; H: (x, z) <- (-z, x); H^-1: (x, z) <- (z, -x); P: z += x; P^-1: z -= x;
  CNOT: x_t += x_c, z_c -= z_t (inverse with the signs flipped); CZ: z_t += x_c,
  z_c += x_t (inverse: -=); SWAP swaps the frames;
* M / M_X (program.py:124-148): M_X first applies the H^-1 rule; the record is
  (reference value + x_q) mod d with the reference's deterministic flag, then z_q
  is drawn again;
* RESET (program.py:150-161): the record is the reference value and flag, x_q = 0,
  z_q drawn again;
* N1 (program.py:163-166): x_q += a, z_q += b with the N1 exponents of that shot.

Randomness: instead of numpy's global generator, Philox4x64-10 keyed (seed,
stream) as everywhere in this build: initial z of qudit q in shot s from counter
(q, s, 1, 0); the redraw after the measurement at instruction k from (k, s, 1, 1);
N1 exponents from the tableau path's draw (k, s, 0, 0) (oracle/philox.py), so a
shot's noise realisation is the same in both paths.  Values are reduced every op
(the reference reduces every 64 ops, identical mod d).
"""
import numpy as np

from . import philox

FRAME_DOMAIN = 1


def frame_draw(seed, stream, c0, shot, kind, d):
    """Uniform value in [0, d) from Philox counter (c0, shot, FRAME_DOMAIN, kind)."""
    out = philox.philox4x64_10((c0, shot, FRAME_DOMAIN, kind), (seed, stream))
    return philox.mulhi64(out[0], d)


def simulate_frame(n, d, ops, ref, shots, seed, stream, noise=None, shot0=1):
    """Frame-sample `shots` shots (global shot indices shot0 .. shot0 + shots - 1).

    ops: IR triples (gate_id, q, t); ref: list of (value, deterministic) per
    measurement event (M, M_X, RESET in op order) of the reference shot;
    noise: {op index: (channel code, keep53)} for N1 ops.
    Returns int64 records [shots, n_events, 5] = (q, round, shot, det, value),
    shot numbered from shot0.
    """
    noise = noise or {}
    ops = [tuple(int(v) for v in op) for op in ops]
    n_ev = sum(1 for g, _, _ in ops if g in (14, 15, 16))
    out = np.zeros((shots, n_ev, 5), dtype=np.int64)
    for s in range(shots):
        shot = shot0 + s
        x = [0] * n
        z = [frame_draw(seed, stream, q, shot, 0, d) for q in range(n)]
        rounds = [0] * n
        ev = 0
        for k, (g, q, t) in enumerate(ops):
            if g == 5:
                x[q], z[q] = (-z[q]) % d, x[q]
            elif g == 6:
                x[q], z[q] = z[q], (-x[q]) % d
            elif g == 7:
                z[q] = (z[q] + x[q]) % d
            elif g == 8:
                z[q] = (z[q] - x[q]) % d
            elif g == 9:
                x[t] = (x[t] + x[q]) % d
                z[q] = (z[q] - z[t]) % d
            elif g == 10:
                x[t] = (x[t] - x[q]) % d
                z[q] = (z[q] + z[t]) % d
            elif g == 11:
                z[t] = (z[t] + x[q]) % d
                z[q] = (z[q] + x[t]) % d
            elif g == 12:
                z[t] = (z[t] - x[q]) % d
                z[q] = (z[q] - x[t]) % d
            elif g == 13:
                x[q], x[t] = x[t], x[q]
                z[q], z[t] = z[t], z[q]
            elif g in (14, 15):
                if g == 15:
                    x[q], z[q] = z[q], (-x[q]) % d
                val, det = ref[ev]
                out[s, ev] = (q, rounds[q], shot, int(det), (val + x[q]) % d)
                rounds[q] += 1
                ev += 1
                z[q] = frame_draw(seed, stream, k, shot, 1, d)
            elif g == 16:
                val, det = ref[ev]
                out[s, ev] = (q, rounds[q], shot, int(det), val)
                rounds[q] += 1
                ev += 1
                x[q] = 0
                z[q] = frame_draw(seed, stream, k, shot, 1, d)
            elif g == 17:
                chan, keep53 = noise[k]
                a, b = philox.noise_exponents(seed, stream, k, shot, d, chan, keep53)
                x[q] = (x[q] + a) % d
                z[q] = (z[q] + b) % d
            # 0 (I) and 1-4 (X, Z and inverses) leave the frame unchanged
    return out
