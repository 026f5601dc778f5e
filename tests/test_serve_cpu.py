"""The serving loop (runtime.cpp serve_loop) on the CPU build of the kernels: requests of
very different cost (batched by cost class, the last in-flight slot kept for small requests)
arrive in bursts and in order; every request's records equal the same circuit run alone.
"""
import numpy as np

from cpu_sim.sim_engine import SimEngine


def test_serving_mixed_costs_equal_alone():
    from sdim_amd import workloads as W
    from sdim_amd.serve import Server
    rng = np.random.default_rng(3)
    reqs = [W.serving_request(rng, k, seed=3, n_lo=4, n_hi=96, shots_hi=300) for k in range(40)]
    eng = SimEngine("plain")
    alone = []
    for r in reqs:
        tk = eng.submit([r]).wait()
        alone.append(tk.records(0).copy())
        tk.release()
    for max_inflight, max_batch in ((4, 8), (1, 64), (2, 3)):
        srv = Server(eng, hbm_budget_bytes=1 << 30, max_inflight=max_inflight, max_batch=max_batch)
        ids = [srv.submit(r) for r in reqs]
        for k, rid in enumerate(ids):
            srv.wait(rid, timeout=600)
            rec, _ = srv.result(rid)
            assert np.array_equal(rec.reshape(alone[k].shape), alone[k]), (max_inflight, max_batch, k)
        assert srv.stats()["completed"] == len(reqs)
        srv.close()
    eng.close()
