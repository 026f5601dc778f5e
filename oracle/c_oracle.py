"""ctypes wrapper of the C restatement (oracle/sdim_oracle.c).  Test infrastructure only."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.oracle_simulate.restype = ctypes.c_int
        L.oracle_simulate.argtypes = [
            ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_void_p,
            ctypes.c_void_p, ctypes.c_void_p,
            ctypes.c_int64, ctypes.c_int64, ctypes.c_uint64, ctypes.c_uint64,
            ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
        _LIB = L
    return _LIB


def simulate(n, d, ops, shots=1, seed=0, stream=0, noise=None, keep_shots=0, shot0=0):
    """Same contract as oracle.program.simulate, computed in C.

    Returns (records int64[shots, n_meas, 5], tableaux list of (x, z, ph, dx, dz, dph) int64).
    """
    ops = np.ascontiguousarray(np.asarray(ops, dtype=np.int64).reshape(-1, 3))
    n_ops = ops.shape[0]
    n_meas = int(np.isin(ops[:, 0], (14, 15, 16)).sum())
    chan = np.zeros(max(n_ops, 1), dtype=np.int32)
    keep = np.zeros(max(n_ops, 1), dtype=np.uint64)
    for k, (c, k53) in (noise or {}).items():
        chan[k] = c
        keep[k] = k53
    rec = np.zeros((shots, max(n_meas, 1), 5), dtype=np.int64)
    tab_words = 4 * n * n + 2 * n
    tab = np.zeros((max(keep_shots, 1), tab_words), dtype=np.int32) if keep_shots else None
    rc = lib().oracle_simulate(
        n, d, n_ops, ops.ctypes.data, chan.ctypes.data, keep.ctypes.data,
        shot0, shots, seed, stream, n_meas, rec.ctypes.data, keep_shots,
        tab.ctypes.data if tab is not None else None)
    if rc != 0:
        raise ValueError(f"oracle_simulate failed: {rc}")
    tabs = []
    if tab is not None:
        nn = n * n
        for s in range(keep_shots):
            b = tab[s].astype(np.int64)
            x = b[0:nn].reshape(n, n)
            z = b[nn:2 * nn].reshape(n, n)
            dx = b[2 * nn:3 * nn].reshape(n, n)
            dz = b[3 * nn:4 * nn].reshape(n, n)
            ph = b[4 * nn:4 * nn + n]
            dph = b[4 * nn + n:4 * nn + 2 * n]
            tabs.append((x, z, ph, dx, dz, dph))
    return rec[:, :n_meas, :], tabs


def frame(n, d, ops, ref, shots, seed=0, stream=0, noise=None, shot0=1):
    """Same contract as oracle.frame.simulate_frame, computed in C.

    ref: (value, deterministic) per measurement event of the reference shot.
    Returns int64 records [shots, n_meas, 5] = (q, round, shot, det, value).
    """
    L = lib()
    if not hasattr(L, "_frame_ready"):
        L.oracle_frame.restype = ctypes.c_int
        L.oracle_frame.argtypes = [
            ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_void_p,
            ctypes.c_void_p, ctypes.c_void_p,
            ctypes.c_int64, ctypes.c_int64, ctypes.c_uint64, ctypes.c_uint64,
            ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
        L._frame_ready = True
    ops = np.ascontiguousarray(np.asarray(ops, dtype=np.int64).reshape(-1, 3))
    n_ops = ops.shape[0]
    n_meas = int(np.isin(ops[:, 0], (14, 15, 16)).sum())
    if len(ref) != n_meas:
        raise ValueError("one reference record per measurement event")
    refb = np.zeros(max(n_meas, 1), dtype=np.uint8)
    for e, (v, det) in enumerate(ref):
        refb[e] = (int(v) & 0x7F) | (0x80 if det else 0)
    chan = np.zeros(max(n_ops, 1), dtype=np.int32)
    keep = np.zeros(max(n_ops, 1), dtype=np.uint64)
    for k, (c, k53) in (noise or {}).items():
        chan[k] = c
        keep[k] = k53
    rec = np.zeros((shots, max(n_meas, 1), 5), dtype=np.int64)
    rc = L.oracle_frame(n, d, n_ops, ops.ctypes.data, chan.ctypes.data, keep.ctypes.data,
                        shot0, shots, seed, stream, n_meas, refb.ctypes.data, rec.ctypes.data)
    if rc != 0:
        raise ValueError(f"oracle_frame failed: {rc}")
    return rec[:, :n_meas, :]
