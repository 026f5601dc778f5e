"""GPU parity: the HIP engine against the oracle, bit-exact, through the C ABI.

Compares measurement records (value, deterministic flag, qudit, round, shot) and
the final tableau of every shot with the C restatement of the oracle
(oracle/sdim_oracle.c, itself pinned by tests/test_oracle_pinning.py) on the
same Philox key.
"""
import numpy as np
import pytest

from parity_helpers import check_job, check_ticket, clifford_job, rand_job

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    from sdim_amd.engine import Engine
    return Engine()


@pytest.mark.parametrize("mode", ["batched", "plain"])
@pytest.mark.parametrize("d", [2, 3, 5, 7, 11])
def test_random_circuits_small(engine, mode, d):
    rng = np.random.default_rng(100 * d + (mode == "plain"))
    jobs = [rand_job(rng, int(rng.integers(1, 12)), d, int(rng.integers(1, 60)), stream=i) for i in range(12)]
    print(f"[parity] mode={mode} d={d} jobs={len(jobs)}", flush=True)
    tk = engine.run(jobs, mode=mode)
    check_ticket(tk, jobs)
    tk.release()


@pytest.mark.parametrize("mode,opts", [("batched", dict()), ("batched", dict(serial_max=8)),
                                       ("batched", dict(allow_lds=0)), ("plain", dict())])
def test_small_class_boundary_and_trivial_circuits(mode, opts):
    """n = 15, 16, 17 around the limit of the small-tableau launch class (with up to 65 shots,
    around its shot limit), an empty circuit, a measurement-only and a noise-only circuit (noise
    probabilities 1, 0 and 1/2), every dimension."""
    from sdim_amd.engine import Engine, Job
    rng = np.random.default_rng(9)
    jobs = []
    for d in (2, 3, 5, 7, 11):
        for n in (15, 16, 17):
            jobs.append(rand_job(rng, n, d, 120, shots=int(rng.integers(1, 66)), stream=len(jobs), same=0.1))
        jobs.append(Job(3, d, np.zeros((0, 3), np.int64), shots=2, seed=1, stream=len(jobs)))
        jobs.append(Job(4, d, np.asarray([(14, q, -1) for q in range(4)] * 2, np.int64), shots=3, seed=2,
                        stream=len(jobs)))
        jobs.append(Job(3, d, np.asarray([(17, q % 3, -1) for q in range(9)], np.int64), shots=5, seed=3,
                        stream=len(jobs), noise_channel=np.asarray([q % 3 for q in range(9)], np.int32),
                        noise_prob=np.asarray([1.0, 0.0, 0.5] * 3)))
    eng = Engine(**opts)
    print(f"[parity-boundary] mode={mode} opts={opts}", flush=True)
    tk = eng.run(jobs, mode=mode)
    check_ticket(tk, jobs)
    tk.release()
    eng.close()


@pytest.mark.parametrize("mode,opts", [("batched", dict()), ("batched", dict(serial_max=1000)),
                                       ("batched", dict(serial_max=0, allow_lds=0)),
                                       ("batched", dict(hbm_threads=64, allow_lds=0, serial_max=8)),
                                       ("plain", dict())])
def test_two_qudit_gates_on_one_qudit(mode, opts):
    """SUM, SUM^-1, CZ, CZ^-1 and SWAP whose two qudits coincide, among random circuits of every
    dimension: serial layers of small tableaux (rows in registers), parallel layers, HBM items,
    the plain path."""
    from sdim_amd.engine import Engine
    eng = Engine(**opts)
    rng = np.random.default_rng(700 + len(opts) + (mode == "plain"))
    jobs = [rand_job(rng, int(rng.integers(1, 20)), d, int(rng.integers(10, 80)), stream=i, same=0.4)
            for i, d in enumerate([2, 3, 5, 7, 11] * 3)]
    print(f"[parity-same-qudit] mode={mode} opts={opts}", flush=True)
    tk = eng.run(jobs, mode=mode)
    check_ticket(tk, jobs)
    tk.release()
    eng.close()


@pytest.mark.parametrize("opts", [dict(), dict(allow_lds=0), dict(shots_per_item=1), dict(threads=256),
                                  dict(threads=512, allow_lds=0), dict(serial_max=0),
                                  dict(serial_max=1000, allow_lds=0), dict(interleave=0, allow_lds=0),
                                  dict(interleave=2, allow_lds=0, lds_bytes=32768)])
def test_medium_circuits_engine_options(opts):
    import zlib
    from sdim_amd.engine import Engine
    eng = Engine(**opts)
    rng = np.random.default_rng(zlib.crc32(repr(sorted(opts.items())).encode()) & 0xFFFF)
    jobs = []
    for i, (n, d) in enumerate([(40, 2), (70, 3), (33, 5), (50, 7), (21, 11), (130, 2), (96, 3)]):
        jobs.append(clifford_job(rng, n, d, 4 * n, shots=3, stream=i, noise_p=0.05 if i % 2 else 0.0))
    print(f"[parity-medium] opts={opts}", flush=True)
    tk = eng.run(jobs)
    check_ticket(tk, jobs)
    tk.release()


@pytest.mark.parametrize("d", [3, 5, 7])
def test_spilling_tableaux(d):
    """n large enough that the odd-d tableau cannot stay in LDS."""
    from sdim_amd.engine import Engine
    eng = Engine()
    rng = np.random.default_rng(7 * d)
    jobs = [clifford_job(rng, 230, d, 2000, shots=2, stream=0)]
    print(f"[parity-spill] d={d}", flush=True)
    tk = eng.run(jobs)
    st = tk.stats()
    assert st["lds_items"] == 0
    check_ticket(tk, jobs)
    tk.release()


def test_replay_recomputes_records():
    """Replays recompute every record: the record arena is poisoned before each replay, and
    afterwards the records match the oracle again and the byte counters are nonzero."""
    from sdim_amd.engine import Engine
    eng = Engine()
    rng = np.random.default_rng(5)
    jobs = [clifford_job(rng, n, d, 3 * n, shots=4, stream=i) for i, (n, d) in enumerate([(30, 2), (60, 3), (220, 5)])]
    tk = eng.run(jobs)
    for which in ("all", "hbm", "lds", "all"):
        tk.poison_records(0xFF)
        assert all((tk.records(i)['measurement_value'] == 127).all() for i in range(len(jobs)))
        tk.replay(2, which)
        st = tk.stats()
        if which == "all":
            assert st["bytes_hbm"] > 0 and st["bytes_lds"] > 0
            check_ticket(tk, jobs, tableaux=False)
        else:
            assert (st["bytes_hbm"] if which == "hbm" else st["bytes_lds"]) > 0
    check_ticket(tk, jobs)
    tk.release()


def test_uneven_batch_all_circuits_digests():
    """Every circuit of the uneven batch (4096 circuits, n ~ U[32,512], d ~ U{2,3,5,7}, 10n ops
    + measure-all, 8 shots) in one submission, then one graph replay: each circuit's records of
    all 8 shots and shot 0's final tableau against the oracle's digests
    (tests/golden/make_uneven_digests.py)."""
    import os
    from parity_helpers import digest_records, digest_tableau, records_array
    from sdim_amd import workloads as W
    from sdim_amd.engine import Engine
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "uneven_batch_digests.npz"))
    wl = W.uneven_batch()
    assert [j.n for j in wl.jobs] == list(g["n"]) and [j.d for j in wl.jobs] == list(g["d"])
    eng = Engine()
    print("[parity-uneven-all] submitting 4096 circuits", flush=True)
    tk = eng.run(wl.jobs)
    for rnd in range(2):
        bad = []
        for i, j in enumerate(wl.jobs):
            if digest_records(records_array(tk.records(i))) != g["records"][i].tobytes():
                bad.append((i, "records"))
            if rnd == 0 and digest_tableau(tk.tableau(i, 0)) != g["tableau0"][i].tobytes():
                bad.append((i, "tableau"))
        print(f"[parity-uneven-all] round {rnd}: {len(bad)} mismatches", flush=True)
        assert not bad, f"{len(bad)} circuits differ from the oracle, first {bad[:8]}"
        if rnd == 0:
            tk.poison_records(0xFF)
            tk.replay(1)
    tk.release()
    eng.close()


def test_list_overflow_fails_ticket_for_good():
    """A kernel list that would overflow (forced from the host with a tiny list_cap) drops the
    write, every kernel exits normally, and the ticket fails: wait, a repeated fetch and a
    replay all report the error."""
    from sdim_amd._lib import EngineError
    from sdim_amd.engine import Engine
    eng = Engine(list_cap=1)
    rng = np.random.default_rng(17)
    jobs = [clifford_job(rng, 40, 3, 200, shots=2, stream=0)]
    tk = eng.submit(jobs)
    for _ in range(2):
        with pytest.raises(EngineError, match="capacity check failed"):
            tk.wait()
    with pytest.raises(EngineError, match="capacity check failed"):
        tk.records(0)
    with pytest.raises(EngineError, match="capacity check failed"):
        tk.replay(1)
    with pytest.raises(EngineError):
        tk.release()
    eng.close()
    # the same circuit without the cap is exact
    eng = Engine()
    tk = eng.run(jobs)
    check_ticket(tk, jobs)
    tk.release()
    eng.close()


def test_uneven_batch_workload_sample():
    """The whole uneven-batch workload (4096 circuits, n ~ U[32,512], d ~ U{2,3,5,7}, 10n ops
    {H,P,SUM,X,Z} + measure-all, 8 shots) in one submission; a sample of its circuits spread
    over n and d is checked against the oracle: records of all 8 shots, tableaux of two."""
    from sdim_amd import workloads as W
    from sdim_amd.engine import Engine
    wl = W.uneven_batch()
    eng = Engine()
    print("[parity-uneven] submitting 4096 circuits", flush=True)
    tk = eng.run(wl.jobs)
    st = tk.stats()
    assert 0 < st["lds_items"] < st["items"]
    by_key = {}
    for i, j in enumerate(wl.jobs):
        by_key.setdefault((j.d, j.n // 64), i)          # first circuit of each (d, n band)
    sample = sorted(by_key.values())
    big = max(range(len(wl.jobs)), key=lambda i: (wl.jobs[i].d != 2, wl.jobs[i].n))
    if big not in sample:
        sample.append(big)
    for i in sample:
        print(f"[parity-uneven] job {i} n={wl.jobs[i].n} d={wl.jobs[i].d}", flush=True)
        check_job(tk, i, wl.jobs[i], tableaux=True, shots=(0, 7))
    tk.release()
    eng.close()


@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_plain_small_configuration(mode):
    """d=2, n=24, 500 uniform {H, P, CNOT} then measure-all, 1024 shots: the records of every
    shot and the final tableaux of four shots against the oracle."""
    from sdim_amd import workloads as W
    from sdim_amd.engine import Engine
    wl = W.plain_small()
    eng = Engine()
    print(f"[parity-config] plain_small mode={mode}", flush=True)
    tk = eng.run(wl.jobs, mode=mode)
    check_job(tk, 0, wl.jobs[0], tableaux=True, shots=(0, 1, 511, 1023))
    tk.release()
    eng.close()


def test_published_random_workload():
    """The reference's published timing workload (100,000 gates CNOT 40% / H 30% / P 30%,
    d in {2,3,5,7}, n in 2..14, one shot, no measurement): the final tableau of all 52
    circuits batched, and of n = 2 and n = 14 per d on the plain path, against the oracle."""
    from sdim_amd import workloads as W
    from sdim_amd.engine import Engine
    from parity_helpers import oracle_of
    wl = W.published_random()
    eng = Engine(serial_max=0)
    print("[parity-config] published_random batched", flush=True)
    tk = eng.run(wl.jobs)
    print("[parity-config] published_random batched, serial layers", flush=True)
    es = Engine(serial_max=1000)
    ts = es.run(wl.jobs)
    sub = [i for i, j in enumerate(wl.jobs) if j.n in (2, 14)]
    print("[parity-config] published_random plain", flush=True)
    tp = eng.run([wl.jobs[i] for i in sub], mode="plain")
    for i, job in enumerate(wl.jobs):
        _, tabs = oracle_of(job, 1)
        for name, a, b in zip(("x", "z", "ph", "dx", "dz", "dph"), tk.tableau(i, 0), tabs[0]):
            assert np.array_equal(a, b), f"tableau {name} differs, circuit {i} (n={job.n}, d={job.d})"
        for name, a, b in zip(("x", "z", "ph", "dx", "dz", "dph"), ts.tableau(i, 0), tabs[0]):
            assert np.array_equal(a, b), f"serial-layer tableau {name} differs, circuit {i} (n={job.n}, d={job.d})"
        if i in sub:
            for name, a, b in zip(("x", "z", "ph", "dx", "dz", "dph"), tp.tableau(sub.index(i), 0), tabs[0]):
                assert np.array_equal(a, b), f"plain tableau {name} differs, circuit {i} (n={job.n}, d={job.d})"
    tk.release()
    tp.release()
    ts.release()
    es.close()
    eng.close()


@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_repetition_configuration(mode):
    """d=3 repetition ring (17 data + 17 ancilla, 10 rounds of SUMs, ancilla M and RESET, X^a /
    Z^b errors with p=1e-3 after every SUM), 4096 shots: every record against the oracle."""
    from sdim_amd import workloads as W
    from sdim_amd.engine import Engine
    wl = W.repetition_code()
    eng = Engine()
    print(f"[parity-config] repetition mode={mode}", flush=True)
    tk = eng.run(wl.jobs, mode=mode)
    check_job(tk, 0, wl.jobs[0], tableaux=True, shots=(0, 4095))
    rec = tk.records(0)
    assert rec.shape == (340, 4096)
    tk.release()
    eng.close()


def test_autotune_keeps_records(tmp_path):
    """The autotuner times candidate launch settings on a small mixed batch; every candidate
    must reproduce the default's records, and the chosen options run bit-exact."""
    from sdim_amd import autotune
    from sdim_amd.engine import Engine
    rng = np.random.default_rng(31)
    jobs = [clifford_job(rng, n, d, 3 * n, shots=4, stream=i) for i, (n, d) in
            enumerate([(40, 2), (90, 3), (230, 5), (300, 7), (20, 11)])]
    space = [("hbm_threads", (512, 256)), ("shots_per_item", (0, 1, 2)), ("lds_bytes", (163584, 65536))]
    out = autotune.tune(jobs, reps=1, space=space, cache=str(tmp_path / "t.json"))
    assert out["ms"] > 0 and all("error" not in t for t in out["trials"])
    assert autotune.cached(jobs, str(tmp_path / "t.json")) == out["options"]
    eng = Engine(**out["options"])
    tk = eng.run(jobs)
    check_ticket(tk, jobs, tableaux=False)
    tk.release()
    eng.close()


def test_chp_files_straight_to_device(tmp_path):
    """.chp files read straight into device gate lists (sdim_amd.circuit_io.read_job, no
    CircuitInstruction objects): the reference's EPR and Steane samples and generated files
    with every gate name, alias and N1 parameters, against the oracle."""
    import os
    from sdim_amd.circuit_io import read_job
    from test_api_cpu import _random_chp_text
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    files = [os.path.join(golden, "epr.chp"), os.path.join(golden, "css_steane_final.chp")]
    rng = np.random.default_rng(99)
    for k, d in enumerate((2, 3, 5, 7, 11)):
        p = tmp_path / f"gen{k}.chp"
        p.write_text(_random_chp_text(rng, int(rng.integers(3, 40)), d, 400))
        files.append(str(p))
    jobs = [read_job(f, shots=6, seed=1234 + k, stream=k) for k, f in enumerate(files)]
    from sdim_amd.engine import Engine
    eng = Engine()
    for mode in ("batched", "plain"):
        print(f"[chp] {len(jobs)} files, {mode}", flush=True)
        tk = eng.run(jobs, mode=mode)
        check_ticket(tk, jobs)
        tk.release()
    eng.close()
