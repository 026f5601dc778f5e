"""Serving: continuous batching of independent circuits on one GPU.

``Server`` wraps the native dispatcher of include/sdim_hip.h (sdim_server_*):
requests are admitted by estimated HBM footprint against a budget, batched
onto free streams while others run, and their latency (arrival to records on
the host) is tracked.  ``poisson_run`` drives it with a Poisson arrival stream.
"""
import ctypes
import time

import numpy as np

from . import _lib
from .engine import MEASUREMENT_DTYPE, Engine


class Server:
    def __init__(self, engine: Engine = None, hbm_budget_bytes=200 << 30, max_inflight=4, max_batch=256):
        self.engine = engine or Engine()
        self.L = self.engine.L
        self.s = ctypes.c_void_p()
        _lib.check(self.L.sdim_server_create(self.engine.ctx, int(hbm_budget_bytes), int(max_inflight),
                                             int(max_batch), ctypes.byref(self.s)), self.L)
        self._jobs = {}

    def submit(self, job):
        rid = ctypes.c_uint64(0)
        cj = job.to_c()
        _lib.check(self.L.sdim_server_submit(self.s, ctypes.byref(cj), ctypes.byref(rid)), self.L)
        self._jobs[rid.value] = job
        return rid.value

    def state(self, rid):
        st = ctypes.c_int32(0)
        _lib.check(self.L.sdim_server_poll(self.s, rid, ctypes.byref(st)), self.L)
        return st.value

    def wait(self, rid, timeout=600.0):
        _lib.check(self.L.sdim_server_wait(self.s, rid, float(timeout)), self.L)

    def result(self, rid):
        """(records as MEASUREMENT_DTYPE (events, shots), latency in seconds); releases the request."""
        nrec = ctypes.c_int64(0)
        lat = ctypes.c_double(0)
        _lib.check(self.L.sdim_server_result(self.s, rid, None, 0, ctypes.byref(nrec), ctypes.byref(lat)), self.L)
        job = self._jobs.pop(rid)
        shots = int(job.shots)
        out = np.zeros(max(nrec.value, 1), dtype=MEASUREMENT_DTYPE)
        _lib.check(self.L.sdim_server_result(self.s, rid, out.ctypes.data, out.shape[0], ctypes.byref(nrec),
                                             ctypes.byref(lat)), self.L)
        return out[:nrec.value].reshape(-1, shots), lat.value

    def stats(self):
        c, p, i = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        gs = ctypes.c_uint64()
        span, p50, p99 = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _lib.check(self.L.sdim_server_stats(self.s, ctypes.byref(c), ctypes.byref(p), ctypes.byref(i),
                                            ctypes.byref(gs), ctypes.byref(span), ctypes.byref(p50),
                                            ctypes.byref(p99)), self.L)
        return {"completed": c.value, "pending": p.value, "inflight": i.value, "gate_shots": gs.value,
                "span_s": span.value, "p50_s": p50.value, "p99_s": p99.value}

    def close(self):
        if self.s:
            self.L.sdim_server_destroy(self.s)
            self.s = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def poisson_run(server, make_request, rate_per_s, n_requests, seed=0):
    """Submit ``n_requests`` requests at Poisson arrival times with mean rate ``rate_per_s``."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_per_s, size=n_requests)
    ids = []
    t0 = time.perf_counter()
    t_next = t0
    for k in range(n_requests):
        t_next += gaps[k]
        delay = t_next - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        ids.append(server.submit(make_request(k)))
    return ids
