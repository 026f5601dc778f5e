"""Serving: continuous batching on the GPU.  Every request's records must be
identical whether it runs alone or in a mix with others (bit-exact, tolerance 0:
integer results), and must match the oracle; admission respects the HBM budget."""
import time

import numpy as np
import pytest

from oracle import c_oracle
from sdim_amd import workloads as W

pytestmark = pytest.mark.gpu


def _requests(k, seed=1, n_hi=160, shots_hi=48):
    rng = np.random.default_rng(seed)
    return [W.serving_request(rng, rid, seed=seed, n_lo=4, n_hi=n_hi, shots_hi=shots_hi) for rid in range(k)]


def test_results_identical_alone_and_in_the_mix():
    from sdim_amd.engine import Engine
    from sdim_amd.serve import Server
    eng = Engine()
    reqs = _requests(40)
    srv = Server(eng, hbm_budget_bytes=8 << 30, max_inflight=4, max_batch=8)
    ids = []
    for k, r in enumerate(reqs):
        ids.append(srv.submit(r))
        if k % 7 == 3:
            time.sleep(0.002)   # arrivals while earlier requests are on the device
    got = []
    for rid in ids:
        srv.wait(rid)
        rec, lat = srv.result(rid)
        assert lat > 0
        got.append(rec)
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(reqs) and st["p99_s"] >= st["p50_s"] > 0
    for r, rec in zip(reqs, got):
        alone = eng.run([r]).records(0)
        assert np.array_equal(alone, rec)
    # a few against the oracle directly
    for r, rec in list(zip(reqs, got))[:6]:
        o, _ = c_oracle.simulate(r.n, r.d, np.asarray(r.ir), shots=r.shots, seed=r.seed, stream=r.stream)
        vals = np.transpose(rec['measurement_value'])
        dets = np.transpose(rec['deterministic']).astype(np.int64)
        assert np.array_equal(vals, o[:, :, 4]) and np.array_equal(dets, o[:, :, 3])


def test_budget_admission():
    from sdim_amd.engine import Engine
    from sdim_amd.serve import Server
    eng = Engine()
    reqs = _requests(12, seed=5)
    # a budget that holds about one request at a time still completes everything
    srv = Server(eng, hbm_budget_bytes=48 << 20, max_inflight=4, max_batch=4)
    ids = [srv.submit(r) for r in reqs]
    for rid in ids:
        srv.wait(rid)
        srv.result(rid)
    assert srv.stats()["completed"] == len(reqs)
    srv.close()
    # a request larger than the whole budget fails cleanly, the server keeps serving
    tiny = Server(eng, hbm_budget_bytes=1 << 20, max_inflight=2, max_batch=2)
    big = W.serving_request(np.random.default_rng(0), 99, n_lo=600, n_hi=700, shots_hi=8)
    rid = tiny.submit(big)
    with pytest.raises(ValueError):
        tiny.wait(rid, timeout=60)
    tiny.close()
