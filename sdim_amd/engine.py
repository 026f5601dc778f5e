"""Python face of the device engine: jobs in, records and tableaux out.

A Job is one circuit in the reference's IR form (``Program._build_ir``'s
(gate_id, qudit_index, target_index) int64 triples, sdim/program.py:512-516)
plus per-op noise parameters, a shot count and a Philox key.
"""
import ctypes
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _lib

# sdim/program.py:34-40
MEASUREMENT_DTYPE = np.dtype([
    ('qudit_index', np.int64),
    ('meas_round', np.int64),
    ('shot', np.int64),
    ('deterministic', np.bool_),
    ('measurement_value', np.int64),
])
assert MEASUREMENT_DTYPE.itemsize == 33

CHANNELS = {"d": 0, "f": 1, "p": 2}


@dataclass
class Job:
    n: int
    d: int
    ir: np.ndarray                       # int64 [n_ops, 3]
    shots: int = 1
    seed: int = 0
    stream: int = 0
    noise_channel: Optional[np.ndarray] = None   # int32 [n_ops]
    noise_prob: Optional[np.ndarray] = None      # float64 [n_ops]
    init: Optional[Sequence[np.ndarray]] = None  # (x, z, ph, dx, dz, dph) reference layout
    _keep: list = field(default_factory=list, repr=False)

    def to_c(self):
        ir = np.ascontiguousarray(np.asarray(self.ir, dtype=np.int64).reshape(-1, 3))
        j = _lib.SdimJob()
        j.n = int(self.n)
        j.d = int(self.d)
        j.n_ops = ir.shape[0]
        j.ir = ir.ctypes.data if ir.size else None
        keep = [ir]
        if self.noise_channel is not None:
            ch = np.ascontiguousarray(self.noise_channel, dtype=np.int32)
            pr = np.ascontiguousarray(self.noise_prob, dtype=np.float64)
            if ch.shape[0] != ir.shape[0] or pr.shape[0] != ir.shape[0]:
                raise ValueError("noise arrays must have one entry per op")
            j.noise_channel = ch.ctypes.data
            j.noise_prob = pr.ctypes.data
            keep += [ch, pr]
        j.shots = int(self.shots)
        j.seed = int(self.seed) & 0xFFFFFFFFFFFFFFFF
        j.stream = int(self.stream) & 0xFFFFFFFFFFFFFFFF
        if self.init is not None:
            arrs = [np.ascontiguousarray(a, dtype=np.int64) for a in self.init]
            for k, a in enumerate(arrs):
                j.init[k] = a.ctypes.data
            keep += arrs
        self._keep = keep
        return j


class Ticket:
    def __init__(self, engine, tid, jobs):
        self.engine = engine
        self.id = tid
        self.jobs = jobs
        self._released = False

    def _check(self, rc):
        _lib.check(rc, self.engine.L)

    def wait(self):
        self._check(self.engine.L.sdim_wait(self.engine.ctx, self.id))
        return self

    def done(self):
        flag = ctypes.c_int32(0)
        self._check(self.engine.L.sdim_poll(self.engine.ctx, self.id, ctypes.byref(flag)))
        return bool(flag.value)

    def num_measurements(self, job=0):
        c = ctypes.c_int64(0)
        self._check(self.engine.L.sdim_num_measurements(self.engine.ctx, self.id, job, ctypes.byref(c)))
        return c.value

    def records(self, job=0):
        """MEASUREMENT_DTYPE array of shape (measurement events, shots)."""
        m = self.num_measurements(job)
        shots = int(self.jobs[job].shots)
        out = np.zeros((m, shots), dtype=MEASUREMENT_DTYPE)
        if m:
            self._check(self.engine.L.sdim_fetch_records(self.engine.ctx, self.id, job, out.ctypes.data, m * shots))
        return out

    def tableau(self, job=0, shot=0):
        """(x, z, ph, dx, dz, dph) int64 in the reference layout."""
        n = int(self.jobs[job].n)
        arrs = [np.zeros((n, n), np.int64), np.zeros((n, n), np.int64), np.zeros(n, np.int64),
                np.zeros((n, n), np.int64), np.zeros((n, n), np.int64), np.zeros(n, np.int64)]
        self._check(self.engine.L.sdim_fetch_tableau(self.engine.ctx, self.id, job, shot,
                                                    *[a.ctypes.data for a in arrs]))
        return tuple(arrs)

    def snapshot(self, job=0, index=0, shot=0):
        """Tableau snapshot `index` of (job, shot): (x, z, ph, dx, dz, dph) in the reference layout
        (tickets submitted with the engine option snapshots=1: per measurement event, or 2: per
        instruction)."""
        n = int(self.jobs[job].n)
        arrs = [np.zeros((n, n), np.int64), np.zeros((n, n), np.int64), np.zeros(n, np.int64),
                np.zeros((n, n), np.int64), np.zeros((n, n), np.int64), np.zeros(n, np.int64)]
        self._check(self.engine.L.sdim_fetch_snapshot(self.engine.ctx, self.id, job, int(index), int(shot),
                                                      *[a.ctypes.data for a in arrs]))
        return tuple(arrs)

    def stats(self):
        ms = ctypes.c_double(0)
        bh = ctypes.c_uint64(0)
        bl = ctypes.c_uint64(0)
        ni = ctypes.c_int64(0)
        nl = ctypes.c_int64(0)
        self._check(self.engine.L.sdim_ticket_stats(self.engine.ctx, self.id, ctypes.byref(ms), ctypes.byref(bh),
                                                   ctypes.byref(bl), ctypes.byref(ni), ctypes.byref(nl)))
        return {"kernel_ms": ms.value, "bytes_hbm": bh.value, "bytes_lds": bl.value,
                "items": ni.value, "lds_items": nl.value}

    def info(self, key):
        """Device layout of the ticket: "items", "lds_items", "gang_items", "gang_members", "gangs", "launches"."""
        v = ctypes.c_int64(0)
        self._check(self.engine.L.sdim_ticket_info(self.engine.ctx, self.id, key.encode(), ctypes.byref(v)))
        return v.value

    def counters(self):
        """Raw device counters: bytes (HBM-resident, LDS-resident) and phase timer ticks."""
        buf = (ctypes.c_uint64 * 18)()
        m = self.engine.L.sdim_ticket_counters(self.engine.ctx, self.id, buf, 18)
        if m < 0:
            self._check(m)
        return [int(buf[i]) for i in range(m)]

    def replay(self, iters=1, which="all"):
        """Re-run the device work `iters` times; returns device milliseconds for all iterations."""
        w = {"all": 0, "hbm": 1, "lds": 2}[which]
        ms = ctypes.c_double(0)
        self._check(self.engine.L.sdim_replay(self.engine.ctx, self.id, int(iters), w, ctypes.byref(ms)))
        return ms.value

    def poison_records(self, value=0xFF):
        """Overwrite the ticket's records (device and host copies) with `value` and clear its
        counters, so that only a replay that does the work restores them (diagnostics)."""
        self._check(self.engine.L.sdim_ticket_poison(self.engine.ctx, self.id, int(value)))

    def release(self):
        if not self._released:
            self._released = True
            self._check(self.engine.L.sdim_release(self.engine.ctx, self.id))

    def __del__(self):
        try:
            if not self._released and self.engine.ctx:
                self.engine.L.sdim_release(self.engine.ctx, self.id)
        except Exception:
            pass


class Engine:
    """One device context.  ``options`` are engine options (see include/sdim_hip.h)."""

    def __init__(self, device=0, **options):
        self.L = _lib.load()
        self.ctx = ctypes.c_void_p()
        _lib.check(self.L.sdim_ctx_create(int(device), ctypes.byref(self.ctx)), self.L)
        for k, v in options.items():
            self.set_option(k, v)

    def set_option(self, key, value):
        _lib.check(self.L.sdim_set_option(self.ctx, key.encode(), int(value)), self.L)

    def get_option(self, key):
        v = ctypes.c_int64(0)
        _lib.check(self.L.sdim_get_option(self.ctx, key.encode(), ctypes.byref(v)), self.L)
        return v.value

    def submit(self, jobs, mode="batched"):
        jobs = list(jobs)
        arr = (_lib.SdimJob * max(len(jobs), 1))()
        for i, j in enumerate(jobs):
            arr[i] = j.to_c()
        tid = ctypes.c_uint64(0)
        m = _lib.MODE_BATCHED if mode == "batched" else _lib.MODE_PLAIN
        _lib.check(self.L.sdim_submit(self.ctx, arr, len(jobs), m, ctypes.byref(tid)), self.L)
        return Ticket(self, tid.value, jobs)

    def run(self, jobs, mode="batched"):
        return self.submit(jobs, mode).wait()

    def frame_submit(self, jobs, refs=None, shot0=None):
        """Pauli-frame sampling (sdim/program.py:45-165, 244-265).

        With ``refs`` (one MEASUREMENT_DTYPE array per job: the reference shot's records, one per
        measurement event in op order) every one of ``jobs[i].shots`` shots, numbered from
        ``shot0`` (default 1), is a frame shot reporting those records shifted by its frame.
        Without ``refs`` the library simulates the reference shot itself (shot 0, N1 as the
        identity, concurrently with the frames) and ``jobs[i].shots`` counts it: records are
        (events, shots) with shot 0 the reference shot, whose final tableau ``tableau(i, 0)``
        returns."""
        jobs = list(jobs)
        arr = (_lib.SdimJob * max(len(jobs), 1))()
        for i, j in enumerate(jobs):
            arr[i] = j.to_c()
        keep = []
        if refs is None:
            ptrs = None
            shot0 = 0 if shot0 is None else shot0
        else:
            if len(refs) != len(jobs):
                raise ValueError("one reference record array per job")
            ptrs = (ctypes.c_void_p * max(len(jobs), 1))()
            for i, r in enumerate(refs):
                r = np.ascontiguousarray(np.asarray(r, dtype=MEASUREMENT_DTYPE).reshape(-1))
                keep.append(r)
                ptrs[i] = r.ctypes.data if r.size else None
            shot0 = 1 if shot0 is None else shot0
        tid = ctypes.c_uint64(0)
        _lib.check(self.L.sdim_frame_submit(self.ctx, arr, len(jobs), ptrs, int(shot0), ctypes.byref(tid)), self.L)
        return Ticket(self, tid.value, jobs)

    def frame_run(self, jobs, refs=None, shot0=None):
        return self.frame_submit(jobs, refs, shot0).wait()

    def close(self):
        if self.ctx:
            self.L.sdim_ctx_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default = None


def default_engine():
    global _default
    if _default is None:
        _default = Engine()
    return _default
