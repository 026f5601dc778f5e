"""Device Pauli-frame sampler (sdim_frame_submit) against the frame oracle
(oracle_frame, pinned by tests/test_frame_cpu.py), bit-exact on the same Philox key:
random circuits of every dimension with all gates, M_X, RESET and N1; ragged shot
blocks; several jobs per submission; frames in LDS and in device memory; the
repetition-code configuration at full size, where every measurement is
deterministic, so the frame shots also equal the tableau path's noisy shots.
"""
import numpy as np
import pytest

from oracle import c_oracle, philox
from parity_helpers import rand_job

pytestmark = pytest.mark.gpu


def _noise(job):
    ir = np.asarray(job.ir).reshape(-1, 3)
    return {k: (int(job.noise_channel[k]), philox.keep_threshold53(float(job.noise_prob[k])))
            for k in range(ir.shape[0]) if ir[k, 0] == 17}


def _noiseless(job):
    from sdim_amd.engine import Job
    ir = np.array(job.ir, dtype=np.int64).reshape(-1, 3).copy()
    ir[ir[:, 0] == 17, 0] = 0
    return Job(job.n, job.d, ir, shots=1, seed=job.seed, stream=job.stream)


def _stack(rec):
    return np.stack([rec['qudit_index'], rec['meas_round'], rec['shot'],
                     rec['deterministic'].astype(np.int64), rec['measurement_value']], axis=-1)


def _reference(eng, jobs):
    tk = eng.run([_noiseless(j) for j in jobs])
    refs = [tk.records(i)[:, 0].copy() for i in range(len(jobs))]
    tk.release()
    return refs


def _check(tk, jobs, refs, shot0):
    for i, (job, ref) in enumerate(zip(jobs, refs)):
        got = np.transpose(_stack(tk.records(i)), (1, 0, 2))          # [shots, events, 5]
        pairs = [(int(r['measurement_value']), bool(r['deterministic'])) for r in ref]
        exp = c_oracle.frame(job.n, job.d, np.asarray(job.ir), pairs, job.shots, job.seed, job.stream,
                             _noise(job), shot0=shot0)
        assert np.array_equal(got, exp), f"frame records differ for job {i} (n={job.n}, d={job.d})"


@pytest.fixture(scope="module")
def eng():
    from sdim_amd.engine import Engine
    return Engine()


@pytest.mark.parametrize("d", [2, 3, 5, 7, 11])
def test_frame_random_circuits(eng, d):
    rng = np.random.default_rng(2000 + d)
    jobs = [rand_job(rng, int(n), d, 8 * int(n), shots=int(s), stream=i)
            for i, (n, s) in enumerate([(1, 5), (7, 300), (40, 257), (130, 64)])]
    refs = _reference(eng, jobs)
    print(f"[frame] random d={d}", flush=True)
    tk = eng.frame_run(jobs, refs, shot0=1)
    assert tk.info("items") == 1 + 2 + 2 + 1
    _check(tk, jobs, refs, 1)
    tk.release()


@pytest.mark.parametrize("d", [2, 3, 11])
def test_frame_two_qudit_gates_on_one_qudit(eng, d):
    """CNOT, CZ (and inverses) and SWAP whose two qudits coincide update the frame as the
    reference's in-place statements do."""
    rng = np.random.default_rng(2200 + d)
    jobs = [rand_job(rng, int(n), d, 8 * int(n), shots=int(s), stream=i, same=0.4)
            for i, (n, s) in enumerate([(1, 9), (6, 200), (30, 65)])]
    refs = _reference(eng, jobs)
    print(f"[frame] same-qudit d={d}", flush=True)
    tk = eng.frame_run(jobs, refs, shot0=1)
    _check(tk, jobs, refs, 1)
    tk.release()


@pytest.mark.parametrize("d", [2, 5])
def test_frame_device_memory_frames(d):
    """Frames in device memory (no LDS budget, and n above the default budget) equal LDS frames."""
    from sdim_amd.engine import Engine
    rng = np.random.default_rng(2100 + d)
    jobs = [rand_job(rng, 60, d, 600, shots=300, stream=0), rand_job(rng, 400, d, 2000, shots=70, stream=1)]
    e0 = Engine(frame_lds_bytes=0)
    refs = _reference(e0, jobs)
    print(f"[frame] device-memory frames d={d}", flush=True)
    tk = e0.frame_run(jobs, refs, shot0=1)
    assert tk.info("lds_items") == 0
    _check(tk, jobs, refs, 1)
    tk.release()
    e0.close()
    e1 = Engine()
    tk = e1.frame_run(jobs, refs, shot0=1)
    assert tk.info("lds_items") == 2          # n = 60 in LDS, n = 400 (100 KB of frames) in device memory
    _check(tk, jobs, refs, 1)
    tk.release()
    e1.close()


def test_frame_shot_offset_and_replay(eng):
    rng = np.random.default_rng(2200)
    jobs = [rand_job(rng, 12, 7, 120, shots=33, stream=5)]
    refs = _reference(eng, jobs)
    tk = eng.frame_run(jobs, refs, shot0=1000)
    _check(tk, jobs, refs, 1000)
    before = tk.records(0).copy()
    tk.replay(3)
    assert np.array_equal(tk.records(0), before)
    with pytest.raises(ValueError):
        tk.tableau(0, 0)
    tk.release()


def test_frame_rejects_bad_reference(eng):
    from sdim_amd.engine import Job
    ir = np.array([(5, 0, -1), (14, 0, -1), (14, 1, -1)], np.int64)
    job = Job(2, 3, ir, shots=4, seed=1)
    ref = _reference(eng, [job])[0]
    bad = ref.copy()
    bad['qudit_index'][1] = 0
    with pytest.raises(ValueError):
        eng.frame_run([job], [bad])
    bad = ref.copy()
    bad['measurement_value'][0] = 3
    with pytest.raises(ValueError):
        eng.frame_run([job], [bad])
    # no measurements at all
    job0 = Job(2, 3, np.array([(5, 0, -1)], np.int64), shots=4, seed=1)
    tk = eng.frame_run([job0], [np.zeros(0, dtype=ref.dtype)])
    assert tk.records(0).shape == (0, 4)
    tk.release()


def test_repetition_configuration_frames(eng):
    """The d=3 repetition-code configuration (17+17 qutrits, 10 rounds, p=1e-3) at 4096
    shots: shot 0 is the noiseless tableau shot, shots 1..4095 frame shots, bit-exact
    against the frame oracle; every measurement is deterministic, so the M records of
    each frame shot equal the tableau path's noisy shot with the same index."""
    from sdim_amd import workloads as W
    from sdim_amd.engine import Job
    job = W.repetition_code().jobs[0]
    refs = _reference(eng, [job])
    assert refs[0]['deterministic'].all()
    fj = Job(job.n, job.d, job.ir, shots=job.shots - 1, seed=job.seed, stream=job.stream,
             noise_channel=job.noise_channel, noise_prob=job.noise_prob)
    print("[frame] repetition configuration", flush=True)
    tk = eng.frame_run([fj], refs, shot0=1)
    _check(tk, [fj], refs, 1)
    fr = tk.records(0)
    tk.release()
    tt = eng.run([job])
    tab = tt.records(0)[:, 1:]
    tt.release()
    ir = np.asarray(job.ir)
    meas = ir[np.isin(ir[:, 0], (14, 15, 16)), 0] == 14
    assert np.array_equal(fr[meas], tab[meas])
    assert (fr['measurement_value'][meas] != 0).any()     # the noise shows


@pytest.mark.parametrize("d", [2, 3, 5, 7, 11])
def test_frame_with_device_reference_shot(eng, d):
    """Without reference records the library runs the reference shot itself (shot 0, N1 as
    the identity) concurrently with the frames: shot 0 equals the noiseless tableau oracle,
    shots 1.. the frame oracle on those records, and tableau(0) the reference final tableau."""
    rng = np.random.default_rng(2300 + d)
    jobs = [rand_job(rng, int(n), d, 8 * int(n), shots=int(s), stream=i)
            for i, (n, s) in enumerate([(5, 1), (9, 300), (60, 70)])]
    print(f"[frame] device reference d={d}", flush=True)
    tk = eng.frame_run(jobs)
    for i, job in enumerate(jobs):
        ir = np.asarray(job.ir)
        ir0 = ir.copy()
        ir0[ir0[:, 0] == 17, 0] = 0
        rec0, tabs = c_oracle.simulate(job.n, d, ir0, shots=1, seed=job.seed, stream=job.stream, keep_shots=1)
        got = np.transpose(_stack(tk.records(i)), (1, 0, 2))              # [shots, events, 5]
        assert np.array_equal(got[0], rec0[0]), f"reference shot differs for job {i}"
        pairs = [(int(r[4]), bool(r[3])) for r in rec0[0]]
        exp = c_oracle.frame(job.n, d, ir, pairs, job.shots - 1, job.seed, job.stream, _noise(job), shot0=1)
        assert np.array_equal(got[1:], exp), f"frame shots differ for job {i}"
        for a, b in zip(tk.tableau(i, 0), tabs[0]):
            assert np.array_equal(a, b)
    before = [tk.records(i).copy() for i in range(len(jobs))]
    tk.replay(2)
    for i in range(len(jobs)):
        assert np.array_equal(tk.records(i), before[i])
    with pytest.raises(ValueError):
        tk.tableau(1, 1)
    tk.release()
