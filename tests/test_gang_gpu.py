"""Gang path: one circuit processed by a gang of workgroups, each owning a range of
generator columns (large tableaux, e.g. d=5 n=8192).  Bit-exact against the C
oracle at sizes the oracle finishes in seconds (the gang size is forced small
with the "gang_min_n" / "gang_chunk" options so that chunk boundaries, empty
members and the cross-member deterministic scan are all exercised), and at the
full large-tableau size through size-independent checks: the oracle on a
shortened circuit of the same shape, and agreement between different gang sizes
and the one-workgroup path.
"""
import numpy as np
import pytest

from parity_helpers import check_ticket, clifford_job, oracle_of, rand_job

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gang_engine():
    from sdim_amd.engine import Engine
    return Engine(gang_min_n=48, gang_chunk=16)


def _records(tk, j):
    rec = tk.records(j)
    return np.stack([rec['qudit_index'], rec['meas_round'], rec['shot'],
                     rec['deterministic'].astype(np.int64), rec['measurement_value']], axis=-1)


@pytest.mark.parametrize("d", [2, 3, 5, 7, 11])
def test_gang_random_circuits(gang_engine, d):
    rng = np.random.default_rng(4000 + d)
    sizes = (130, 300, 620) if d == 2 else (60, 150, 260)
    jobs = [rand_job(rng, n, d, 3 * n, shots=3, stream=i) for i, n in enumerate(sizes)]
    print(f"[gang] random d={d} n={sizes}", flush=True)
    tk = gang_engine.run(jobs)
    assert tk.info("gang_items") == len(jobs)
    assert tk.info("gang_members") >= 2
    check_ticket(tk, jobs)
    tk.release()


@pytest.mark.parametrize("d", [2, 3, 5, 7, 11])
def test_gang_measure_all_with_noise(gang_engine, d):
    rng = np.random.default_rng(5000 + d)
    n = 520 if d == 2 else 210
    jobs = [clifford_job(rng, n, d, 3 * n, shots=5, stream=0, meas_rounds=2, noise_p=0.05)]
    print(f"[gang] measure-all d={d} n={n}", flush=True)
    tk = gang_engine.run(jobs)
    assert tk.info("gang_items") == 1
    check_ticket(tk, jobs)
    tk.release()


@pytest.mark.parametrize("d", [2, 5])
def test_gang_sizes_agree_with_one_workgroup(d):
    """The same circuit on gangs of 2, 8 and 64 members (and of 2 members of 256 threads with
    serial gate layers) and on one workgroup."""
    from sdim_amd.engine import Engine
    rng = np.random.default_rng(6000 + d)
    n = 1100
    job = clifford_job(rng, n, d, 2 * n, shots=4, stream=3, noise_p=0.02)
    rec_o, _ = oracle_of(job, 0)
    runs = []
    for opts in (dict(gang_min_n=0), dict(gang_min_n=256, gang_chunk=16, gang_max=2),
                 dict(gang_min_n=256, gang_chunk=16, gang_max=8), dict(gang_min_n=256, gang_chunk=16, gang_max=64),
                 # serial layers; for d = 5 a member owns more column words than it has threads
                 dict(gang_min_n=256, gang_chunk=16, gang_max=2, gang_threads=256, serial_max=1000),
                 # the LDS cap of the one-workgroup classes leaves gang members their whole LDS
                 dict(gang_min_n=256, gang_chunk=16, gang_max=8, lds_bytes=32768)):
        print(f"[gang] sizes d={d} opts={opts}", flush=True)
        eng = Engine(**opts)
        tk = eng.run([job])
        if opts["gang_min_n"]:
            assert tk.info("gang_members") == min(opts["gang_max"], 16 if d == 2 else 128)
        else:
            assert tk.info("gang_items") == 0
        rec = _records(tk, 0)
        assert np.array_equal(np.transpose(rec, (1, 0, 2)), rec_o), f"records differ from the oracle, opts={opts}"
        runs.append((rec, [tk.tableau(0, s) for s in (0, 3)]))
        tk.release()
        eng.close()
    for rec, tabs in runs[1:]:
        assert np.array_equal(rec, runs[0][0])
        for a, b in zip(tabs, runs[0][1]):
            for u, v in zip(a, b):
                assert np.array_equal(u, v)


def test_gang_initial_tableau_and_replay(gang_engine):
    """A given initial tableau on a gang equals the one-workgroup result; replays reproduce it."""
    from sdim_amd.engine import Engine, Job
    rng = np.random.default_rng(77)
    n, d = 180, 7
    prep = clifford_job(rng, n, d, 2 * n, shots=1, stream=0, meas_rounds=0)
    tk = gang_engine.run([prep])
    init = tuple(np.asarray(a) for a in tk.tableau(0, 0))
    tk.release()
    base = clifford_job(rng, n, d, 2 * n, shots=3, stream=1)
    job = Job(n, d, base.ir, shots=3, seed=base.seed, stream=1, init=init)
    one = Engine(gang_min_n=0)
    ref = one.run([job])
    got = gang_engine.run([job])
    assert got.info("gang_items") == 1
    assert np.array_equal(_records(got, 0), _records(ref, 0))
    for s in range(3):
        for u, v in zip(got.tableau(0, s), ref.tableau(0, s)):
            assert np.array_equal(u, v)
    got.release()
    ref.release()
    one.close()
    # replays of a gang launch reproduce the records
    tk = gang_engine.run([base])
    before = _records(tk, 0).copy()
    tk.replay(3, "all")
    assert np.array_equal(_records(tk, 0), before)
    check_ticket(tk, [base])
    tk.release()


def _large_shape_job(n, d, shots, layers, meas_per_layer, depth_factor, seed):
    from sdim_amd import workloads as W
    from sdim_amd.engine import Job
    rng = np.random.default_rng(seed)
    parts = [W.random_clifford_ir(rng, n, d, depth_factor * n, [5, 7, 9], measure_rounds=0)]
    for _ in range(layers):
        parts.append(W.random_clifford_ir(rng, n, d, n // 2, [9], measure_rounds=0))
        qs = rng.choice(n, meas_per_layer, replace=False)
        parts.append(np.stack([np.full(meas_per_layer, 14), qs, np.full(meas_per_layer, -1)], axis=1))
    ir = np.concatenate(parts).astype(np.int64)
    return Job(n, d, ir, shots=shots, seed=seed, stream=0)


def test_large_tableau_shape_against_oracle():
    """d=5, n=8192 (the large-tableau shape) with a shortened circuit: 2n gates and
    2 x (n/2 SUM + 16 M), 2 shots, records bit-exact against the C oracle."""
    from sdim_amd.engine import Engine
    job = _large_shape_job(8192, 5, shots=2, layers=2, meas_per_layer=16, depth_factor=2, seed=11)
    eng = Engine()
    print("[gang] large shape n=8192 vs oracle", flush=True)
    tk = eng.run([job])
    assert tk.info("gang_items") == 1 and tk.info("gang_members") >= 64
    rec = np.transpose(_records(tk, 0), (1, 0, 2))
    tk.release()
    rec_o, _ = oracle_of(job, 0)
    assert np.array_equal(rec, rec_o)


def test_large_tableau_config_gang_sizes_agree():
    """The full large-tableau configuration (d=5, n=8192, 4n gates, 16 x (n/2 SUM + 64 M),
    32 shots) on the default gang and on gangs of 32: identical records for every shot."""
    from sdim_amd import workloads as W
    from sdim_amd.engine import Engine
    wl = W.large_tableau()
    out = []
    for opts in (dict(), dict(gang_max=32)):
        print(f"[gang] large config opts={opts}", flush=True)
        eng = Engine(**opts)
        tk = eng.run(wl.jobs)
        assert tk.info("gang_items") == 1
        out.append(_records(tk, 0))
        tk.release()
        eng.close()
    assert out[0].shape == (16 * 64, 32, 5)
    assert np.array_equal(out[0], out[1])
    # X/Z blocks are shared by the shots, so whether an outcome is random does not depend on the shot
    assert np.all(out[0][:, :, 3] == out[0][:, :1, 3])


def test_gang_timeout_fails_ticket_and_server_requests():
    """A gang whose barrier wait bound is forced to one tick gives up at its first barrier
    (every member sees the error word and leaves the kernel normally).  The ticket then fails
    for good (wait, a repeated fetch, a replay), and a serving request that ran on such a gang
    reports failure (state 3) instead of records."""
    from sdim_amd._lib import EngineError
    from sdim_amd.engine import Engine
    from sdim_amd.serve import Server
    rng = np.random.default_rng(77)
    jobs = [clifford_job(rng, 200, 3, 600, shots=2, stream=0, meas_rounds=2)]
    eng = Engine(gang_min_n=48, gang_chunk=16, gang_timeout=1)
    print("[gang] forced barrier timeout", flush=True)
    tk = eng.submit(jobs)
    for _ in range(2):
        with pytest.raises(EngineError, match="gang barrier timed out"):
            tk.wait()
    with pytest.raises(EngineError, match="gang barrier timed out"):
        tk.records(0)
    with pytest.raises(EngineError, match="gang barrier timed out"):
        tk.replay(1)
    with pytest.raises(EngineError):
        tk.release()
    srv = Server(eng, hbm_budget_bytes=1 << 34)
    rid = srv.submit(jobs[0])
    with pytest.raises((ValueError, EngineError), match="gang barrier timed out"):
        srv.wait(rid, timeout=120.0)
    assert srv.state(rid) == 3
    srv.close()
    eng.close()
    # the same circuit with the default bound is exact
    eng = Engine(gang_min_n=48, gang_chunk=16)
    tk = eng.run(jobs)
    assert tk.info("gang_items") == 1
    check_ticket(tk, jobs)
    tk.release()
    eng.close()
