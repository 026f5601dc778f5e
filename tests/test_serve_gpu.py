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


def test_serving_mix_at_full_distribution():
    """Requests from the serving configuration's own distribution (n log-uniform in [8, 2048],
    d in {2,3,5,7,11}, depth 8n, shots log-uniform in [1, 4096]) plus forced corners: a
    request on the gang path (n = 2048, d = 11), 4096-shot requests and d = 11.  Each request's
    records are identical alone and in the mix; the oracle checks the first shots of every
    request it finishes quickly; the gang request also equals the one-workgroup path."""
    from sdim_amd.engine import Engine, Job
    from sdim_amd.serve import Server
    rng = np.random.default_rng(2025)
    reqs = [W.serving_request(rng, rid, seed=11) for rid in range(24)]
    corner = np.random.default_rng(7)
    for n, d, shots in ((2048, 11, 8), (512, 7, 4096), (8, 11, 4096), (1600, 2, 16)):
        ir = W.random_clifford_ir(corner, n, d, 8 * n, [W.H, W.P, W.CNOT, W.X, W.Z])
        reqs.append(Job(n, d, ir, shots=shots, seed=11, stream=100 + len(reqs)))
    eng = Engine()
    srv = Server(eng, hbm_budget_bytes=200 << 30, max_inflight=4, max_batch=16)
    print(f"[serve] {len(reqs)} requests, n up to {max(r.n for r in reqs)}, shots up to {max(r.shots for r in reqs)}",
          flush=True)
    ids = []
    for k, r in enumerate(reqs):
        ids.append(srv.submit(r))
        if k % 5 == 2:
            time.sleep(0.003)   # arrivals while earlier requests are on the device
    got = []
    for rid in ids:
        srv.wait(rid, timeout=300)
        got.append(srv.result(rid)[0])
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(reqs)
    for k, (r, rec) in enumerate(zip(reqs, got)):
        tk = eng.run([r])
        if r.n >= 1536:
            assert tk.info("gang_items") == 1, "the large request should run on a gang"
        assert np.array_equal(tk.records(0), rec), f"request {k} (n={r.n}, d={r.d}, shots={r.shots}) differs alone"
        tk.release()
        if r.n <= 160:
            s = min(r.shots, 16)
            o, _ = c_oracle.simulate(r.n, r.d, np.asarray(r.ir), shots=s, seed=r.seed, stream=r.stream)
            vals = np.transpose(rec['measurement_value'][:, :s])
            dets = np.transpose(rec['deterministic'][:, :s]).astype(np.int64)
            assert np.array_equal(vals, o[:, :, 4]) and np.array_equal(dets, o[:, :, 3]), f"request {k} vs oracle"
    big = next(k for k, r in enumerate(reqs) if r.n == 2048)
    r = reqs[big]
    o, _ = c_oracle.simulate(r.n, r.d, np.asarray(r.ir), shots=1, seed=r.seed, stream=r.stream)
    assert np.array_equal(got[big]['measurement_value'][:, 0], o[0, :, 4])
    assert np.array_equal(got[big]['deterministic'][:, 0].astype(np.int64), o[0, :, 3])
    one_wg = Engine(gang_min_n=0)
    tk = one_wg.run([reqs[big]])
    assert tk.info("gang_items") == 0 and np.array_equal(tk.records(0), got[big])
    tk.release()
    one_wg.close()
    eng.close()
