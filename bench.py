#!/usr/bin/env python3
"""Benchmark of the prime-d stabilizer tableau engine on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload uneven|small|repetition]

A step is one pass of the hot path over the whole workload: every circuit of
the batch simulated for all its shots (records and final tableaux written),
with every input already resident in HBM (the captured hipGraph of the batched
path is replayed).  Prints ONE JSON line.

Besides the headline (batched persistent path) the line carries:
  plain          the one-launch-per-primitive-op path on a deterministic subset of
                 the circuits (every k-th circuit), and the batched path on the same subset;
                 speedup_vs_plain is batched over plain on that subset
  per_shot_tableaux  the headline workload with a separate tableau per shot
                 (shots_per_item=1); the headline shares each circuit's X/Z blocks
                 between its shots (config.xz_shared_by_shots)
  roofline       HBM roofline of the kernel pass over HBM-resident tableaux
  roofline_lds   LDS roofline of the kernel pass over LDS-resident tableaux
  large_tableau  the large-tableau configuration (d=5, n=8192, one circuit on a gang
                 of workgroups) with its HBM roofline
  frame          the reference's default shots > 1 path on the repetition-code
                 configuration: a noiseless tableau reference shot + 4095 Pauli-frame
                 shots on the device, alone and as 64 circuits per launch, with the
                 frame oracle (C, 1 core) beside it
  published_workload  the reference's published 100,000-gate circuits (d in {2,3,5,7},
                 n in 2..14, one shot): all 52 batched, and n = 2 / 14 alone beside the
                 published seconds
  serving        Poisson arrivals with HBM-budget admission: p50/p99 latency, aggregate rate
                 (options tuned for the headline workload apply to it only; the other legs
                 run with the engine's defaults)
  cpu_baseline   the CPU oracle (literal restatement of the reference's
                 Program.simulate tableau path, same interpreted cost structure),
                 timed on a bounded sample on this host, 1 core
  cpu_baseline_parallel  the same oracle in one process per host core (up to 16),
                 circuits dealt round-robin, timed before the GPU is touched
"""
import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = "gate-shots/s, batched vs plain on 1 MI355X (bit-exact); p50/p99 latency"
HBM_PEAK_GBS = 8000.0
# 256 CUs x 128 B/clk x 2.4 GHz
LDS_PEAK_GBS = 256 * 128 * 2.4
# full-rate wave64 VALU instructions: 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles
VALU_PEAK = 256 * 4 * 2.4e9 / 4


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="uneven", choices=["uneven", "small", "repetition", "frame"],
                    help="frame: only the Pauli-frame leg (repetition-code configuration)")
    ap.add_argument("--plain-every", type=int, default=16, help="plain path runs every k-th circuit")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--shots-per-item", type=int, default=0)
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-published", action="store_true",
                    help="skip the leg on the reference's published 100,000-gate workload")
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--no-frame", action="store_true")
    ap.add_argument("--no-per-shot", action="store_true",
                    help="skip the leg with a separate tableau per shot (shots_per_item=1)")
    ap.add_argument("--tune", action="store_true",
                    help="run the autotuner on the workload first and measure with its choice")
    ap.add_argument("--defaults", action="store_true",
                    help="use the engine's default options instead of the tuned ones shipped for the workload")
    ap.add_argument("--serve-requests", type=int, default=400)
    ap.add_argument("--serve-rate", type=float, default=0.0, help="requests/s (0: 60%% of a measured burst rate)")
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain elementwise copy of 1 GiB (reads 1 GiB, writes 1 GiB) to calibrate "
                         "rocprof FETCH_SIZE / WRITE_SIZE")
    return ap.parse_args()


def profiled_value(name, key):
    """A figure recorded from a rocprofv3 run in profiles/<name>.md ("key: value"), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", name + ".md")
    try:
        for line in open(path):
            if line.startswith(key + ":"):
                return int(float(line.split(":", 1)[1]))
    except OSError:
        pass
    return None


def profiled_traffic(name):
    """HBM bytes per launch measured by rocprofv3 (FETCH_SIZE + WRITE_SIZE, calibrated against a
    plain copy) in profiles/<name>.md, or None."""
    return profiled_value(name, "traffic_bytes_per_launch")


def copy_calibration():
    import torch
    nbytes = 1 << 30
    x = torch.ones(nbytes // 4, dtype=torch.int32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        torch.bitwise_or(x, 0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"bytes_read_per_launch": nbytes, "bytes_written_per_launch": nbytes,
                                           "launches": 3, "kernel": "torch bitwise_or(int32, 0)"}}), flush=True)


def make_workload(name):
    from sdim_amd import workloads as W
    if name == "uneven":
        return W.uneven_batch()
    if name == "small":
        return W.plain_small()
    return W.repetition_code()


def cpu_baseline(jobs, budget_s):
    """Literal CPU oracle (reference cost structure) on the first circuits of the workload, 1 shot each."""
    from oracle import philox
    from oracle import program as OP
    done_ops, done_jobs, t_total = 0, 0, 0.0
    for j in jobs:
        ir = np.asarray(j.ir).reshape(-1, 3)
        noise = {}
        if j.noise_channel is not None:
            for k in np.nonzero(ir[:, 0] == 17)[0]:
                noise[int(k)] = (int(j.noise_channel[k]), philox.keep_threshold53(float(j.noise_prob[k])))
        t0 = time.perf_counter()
        OP.simulate(j.n, j.d, [tuple(r) for r in ir.tolist()], shots=1, seed=j.seed, stream=j.stream,
                    noise=noise, keep_tableaux=False)
        t_total += time.perf_counter() - t0
        done_ops += int((ir[:, 0] != 0).sum())
        done_jobs += 1
        if t_total >= budget_s:
            break
    ns = [int(j.n) for j in jobs[:done_jobs]]
    return {"value": done_ops / t_total, "unit": "gate-shots/s", "cores": 1, "kind": "port",
            "sample": f"oracle/program.py (literal restatement of sdim Program.simulate tableau path, "
                      f"interpreted per-gate loops) on the first {done_jobs} circuit(s) of the workload "
                      f"(n={ns}), 1 shot each, {done_ops} gate-shots in {t_total:.1f} s",
            "host_cores_available": len(os.sched_getaffinity(0))}


def _cpu_worker(task):
    """One process of the parallel CPU baseline: the literal oracle on its circuits, 1 shot each,
    until its time budget is used (the circuit in progress is finished)."""
    items, budget_s = task
    from oracle import program as OP
    ops_done, t_total = 0, 0.0
    for n, d, ir, seed, stream, noise in items:
        t0 = time.perf_counter()
        OP.simulate(n, d, [tuple(r) for r in ir], shots=1, seed=seed, stream=stream, noise=noise, keep_tableaux=False)
        t_total += time.perf_counter() - t0
        ops_done += sum(1 for r in ir if r[0] != 0)
        if t_total >= budget_s:
            break
    return ops_done, t_total


def cpu_baseline_parallel(jobs, budget_s, procs):
    """The same literal oracle in `procs` processes on the host cores, circuits dealt round-robin
    (circuits and shots are independent, so the reference's work parallelises this way).  Runs
    before this process touches the GPU: the workers are started with the spawn method."""
    import multiprocessing as mp
    from oracle import philox
    tasks = [[] for _ in range(procs)]
    for i, j in enumerate(jobs[:procs * 8]):
        ir = np.asarray(j.ir).reshape(-1, 3)
        noise = {}
        if j.noise_channel is not None:
            for k in np.nonzero(ir[:, 0] == 17)[0]:
                noise[int(k)] = (int(j.noise_channel[k]), philox.keep_threshold53(float(j.noise_prob[k])))
        tasks[i % procs].append((int(j.n), int(j.d), ir.tolist(), int(j.seed), int(j.stream), noise))
    t0 = time.perf_counter()
    with mp.get_context("spawn").Pool(procs) as pool:
        res = pool.map(_cpu_worker, [(t, budget_s) for t in tasks])
    wall = time.perf_counter() - t0
    ops = sum(r[0] for r in res)
    # the processes run concurrently: the aggregate rate is the sum of their rates
    rate = sum(r[0] / r[1] for r in res if r[1] > 0)
    return {"value": rate, "unit": "gate-shots/s", "cores": procs, "kind": "port",
            "sample": f"oracle/program.py in {procs} concurrent processes, circuits of the workload dealt "
                      f"round-robin, 1 shot each; {ops} gate-shots, each process timed over its own circuits "
                      f"({wall:.1f} s wall), rates summed",
            "host_cores_available": len(os.sched_getaffinity(0))}


def serving(eng, n_requests, rate, budget_gb, seed=5):
    """Serving mix: Poisson arrivals of random circuits, n log-uniform in [8, 2048], d in {2,3,5,7,11},
    depth 8n, shots log-uniform in [1, 4096]; admission against an HBM budget; continuous batching."""
    from sdim_amd import workloads as W
    from sdim_amd.serve import Server, poisson_run
    rng = np.random.default_rng(seed)
    reqs = [W.serving_request(rng, k, seed=seed) for k in range(n_requests)]
    gate_shots = W.count_gate_shots(reqs)
    # burst capacity: the first requests submitted at once, completion rate measured (also the
    # calibration of the Poisson rate, 60% of it, when no rate is given)
    srv = Server(eng, hbm_budget_bytes=int(budget_gb * (1 << 30)))
    burst = reqs[:min(100, n_requests)]
    t0 = time.perf_counter()
    ids = [srv.submit(r) for r in burst]
    for rid in ids:
        srv.wait(rid)
        srv.result(rid)
    burst_s = time.perf_counter() - t0
    srv.close()
    burst_gs = W.count_gate_shots(burst)
    if rate <= 0:
        rate = 0.6 * len(burst) / burst_s
    # where the gate-shots are, by size band: a request below gang_min_n runs on one workgroup
    # (LDS-resident while its tableau fits), one at or above it on a gang of workgroups
    def share(lo, hi):
        return round(sum(W.count_gate_shots([r]) for r in reqs if lo <= r.n < hi) / gate_shots, 4)
    gmin = eng.get_option("gang_min_n")
    srv = Server(eng, hbm_budget_bytes=int(budget_gb * (1 << 30)))
    ids = poisson_run(srv, lambda k: reqs[k], rate, n_requests, seed=seed)
    for rid in ids:
        srv.wait(rid)
        srv.result(rid)
    st = srv.stats()
    srv.close()
    return {"requests": n_requests, "arrival_rate_per_s": rate, "p50_latency_s": st["p50_s"],
            "p99_latency_s": st["p99_s"], "gate_shots": gate_shots,
            "aggregate_gate_shots_per_s": st["gate_shots"] / st["span_s"] if st["span_s"] else None,
            "span_s": st["span_s"], "hbm_budget_gb": budget_gb,
            "burst": {"requests": len(burst), "seconds": round(burst_s, 4),
                      "requests_per_s": round(len(burst) / burst_s, 2),
                      "gate_shots_per_s": burst_gs / burst_s},
            "gate_shot_share": {"n<256": share(0, 256), f"256<=n<{gmin} (one-workgroup items)": share(256, gmin),
                                f"n>={gmin} (gang)": share(gmin, 1 << 40)},
            "requests_by_band": [sum(1 for r in reqs if r.n < 256), sum(1 for r in reqs if 256 <= r.n < gmin),
                                 sum(1 for r in reqs if r.n >= gmin)],
            "workload": "Poisson arrivals; n log-U[8,2048], d in {2,3,5,7,11}, depth 8n, shots log-U[1,4096]"}


# wall seconds the reference publishes for one 100,000-gate circuit, one shot
# (tests/testdata/simulation_data.json:23-26, tableau_times at n = 2 and n = 14)
PUBLISHED_S = {(2, 2): 0.13554, (2, 14): 0.60069, (3, 2): 0.13753, (3, 14): 0.60742,
               (5, 2): 0.13361, (5, 14): 0.61090, (7, 2): 0.13227, (7, 14): 0.61197}


def published_workload(eng, steps, cpu_seconds):
    """The reference's published timing workload: d in {2,3,5,7} x n in 2..14, one circuit of
    100,000 gates (CNOT 40% / H 30% / P 30%) each, one shot, no measurement.  All 52 circuits
    batched in one submission; n = 2 and n = 14 per d alone (batched and plain) beside the
    published seconds; the CPU port on one circuit."""
    from sdim_amd import workloads as W
    wl = W.published_random()
    tk = eng.submit(wl.jobs, "batched").wait()
    tk.replay(1)
    t0 = time.perf_counter()
    tk.replay(steps)
    t_all = (time.perf_counter() - t0) / steps
    tk.release()
    rows = []
    for (d, n), pub in PUBLISHED_S.items():
        job = next(j for j in wl.jobs if j.d == d and j.n == n)
        tb = eng.submit([job], "batched").wait()
        tb.replay(1)
        t0 = time.perf_counter()
        tb.replay(steps)
        b_s = (time.perf_counter() - t0) / steps
        tb.release()
        tp = eng.submit([job], "plain").wait()
        t0 = time.perf_counter()
        tp.replay(1)
        p_s = time.perf_counter() - t0
        tp.release()
        rows.append({"d": d, "n": n, "published_s": pub, "batched_alone_s": round(b_s, 5),
                     "plain_s": round(p_s, 4), "batched_vs_published": round(pub / b_s, 1)})
    out = {"workload": wl.name, "description": wl.description, "circuits": len(wl.jobs),
           "gate_shots": wl.gate_shots, "value": wl.gate_shots / t_all, "unit": "gate-shots/s",
           "ms_per_step": 1e3 * t_all, "per_circuit": rows}
    if cpu_seconds > 0:
        from oracle import program as OP
        job = next(j for j in wl.jobs if j.d == 2 and j.n == 14)
        ops = [tuple(r) for r in np.asarray(job.ir).tolist()]
        t0 = time.perf_counter()
        OP.simulate(job.n, job.d, ops, shots=1, seed=job.seed, stream=job.stream, keep_tableaux=False)
        c_s = time.perf_counter() - t0
        out["cpu_port"] = {"seconds": round(c_s, 3), "d": 2, "n": 14, "cores": 1, "kind": "port",
                           "sample": "oracle/program.py on the d=2, n=14 circuit, 1 shot"}
    return out


def large_tableau(eng, steps, warmup):
    """d=5, n=8192: 4n {H,P,SUM} + 16 x (n/2 SUM + 64 M), 32 shots; one item on a gang of workgroups."""
    from sdim_amd import workloads as W
    wl = W.large_tableau()
    tk = eng.submit(wl.jobs, "batched").wait()
    st = tk.stats()
    if warmup:
        tk.replay(warmup)
    t0 = time.perf_counter()
    ms = tk.replay(steps) / steps
    wall = (time.perf_counter() - t0) / steps
    st2 = tk.stats()
    members = tk.info("gang_members")
    tk.release()
    ach = st2["bytes_hbm"] / (ms * 1e-3) / 1e9
    return {"workload": wl.name, "description": wl.description, "gate_shots": wl.gate_shots,
            "value": wl.gate_shots / wall, "unit": "gate-shots/s", "ms_per_step": 1e3 * wall,
            "device_ms_per_step": ms, "gang_members": members,
            "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(ach / HBM_PEAK_GBS, 4),
                         "traffic": profiled_traffic("gang_kernel_large_tableau"),
                         "traffic_source": "profiles/gang_kernel_large_tableau.md",
                         "kernel": "sdim::gang_kernel<5, %d>" % eng.get_option("gang_threads"),
                         "algorithmic_bytes": st2["bytes_hbm"], "kernel_ms": round(ms, 3)}}


def frame_sampler(eng, steps, warmup, cpu_seconds):
    """Repetition-code configuration (d=3, 17+17 qutrits, 10 rounds, p=1e-3, 4096 shots) through
    the reference's default shots > 1 path (sdim/program.py:244-265): shot 0 is a noiseless
    tableau shot, shots 1..4095 are Pauli-frame shots (frame_kernel), both in one submission
    (the reference shot runs concurrently with the frames).  `value` counts all 4096 shots over
    that submission's device time; the two launches are also timed on their own; `batch64`
    runs 64 such circuits (Philox streams 0..63) per submission."""
    from sdim_amd import workloads as W
    from sdim_amd.engine import Job
    job = W.repetition_code().jobs[0]
    ops_per_shot = int((np.asarray(job.ir)[:, 0] != 0).sum())

    def noiseless(j):
        ir = np.array(j.ir).copy()
        ir[ir[:, 0] == 17, 0] = 0
        return Job(j.n, j.d, ir, shots=1, seed=j.seed, stream=j.stream)

    def frames(j):
        return Job(j.n, j.d, j.ir, shots=j.shots - 1, seed=j.seed, stream=j.stream,
                   noise_channel=j.noise_channel, noise_prob=j.noise_prob)

    out = {"workload": "repetition_d3_ring17", "gate_shots": ops_per_shot * job.shots}

    def timed(tk):
        if warmup:
            tk.replay(warmup)
        ms = tk.replay(steps) / steps
        return ms

    # the two launches on their own: the reference shot (tableau) and the frame shots
    tr = eng.submit([noiseless(job)], "batched").wait()
    ref = tr.records(0)[:, 0].copy()
    tf = eng.frame_submit([frames(job)], [ref], shot0=1).wait()
    ms_ref, ms_fr = timed(tr), timed(tf)
    tr.release()
    tf.release()
    # one submission: the library runs the reference shot concurrently with the frames
    tc = eng.frame_submit([job]).wait()
    ms_c = timed(tc)
    tc.release()
    out.update({"value": ops_per_shot * job.shots / (ms_c * 1e-3), "unit": "gate-shots/s", "device_ms": ms_c,
                "reference_shot_ms": ms_ref, "frame_ms": ms_fr,
                "kernels": "sdim::batch_kernel (reference shot) || sdim::frame_kernel<256> (frame shots)"})
    # 64 circuits in one submission
    jobs = [Job(job.n, job.d, job.ir, shots=job.shots, seed=job.seed, stream=s,
                noise_channel=job.noise_channel, noise_prob=job.noise_prob) for s in range(64)]
    tr = eng.submit([noiseless(j) for j in jobs], "batched").wait()
    refs = [tr.records(i)[:, 0].copy() for i in range(64)]
    tf = eng.frame_submit([frames(j) for j in jobs], refs, shot0=1).wait()
    ms_ref, ms_fr = timed(tr), timed(tf)
    tr.release()
    tf.release()
    tc = eng.frame_submit(jobs).wait()
    ms_c = timed(tc)
    tc.release()
    draws = 64 * (job.shots - 1) * (job.n + int(np.isin(np.asarray(job.ir)[:, 0], (14, 15, 16, 17)).sum()))
    insts = profiled_value("frame_kernel_repetition", "valu_insts_per_launch_batch64")
    out["batch64"] = {"value": 64 * ops_per_shot * job.shots / (ms_c * 1e-3), "unit": "gate-shots/s",
                      "device_ms": ms_c, "reference_shots_ms": ms_ref, "frame_ms": ms_fr,
                      "philox_blocks_per_s": draws / (ms_fr * 1e-3),
                      "roofline": None if insts is None else {
                          "bound": "valu", "achieved": insts / (ms_fr * 1e-3), "peak": VALU_PEAK,
                          "unit": "wave64 VALU instructions/s", "frac": round(insts / (ms_fr * 1e-3) / VALU_PEAK, 4),
                          "traffic": None, "source": "profiles/frame_kernel_repetition.md (SQ_INSTS_VALU)"}}
    # the frame oracle (C restatement, one core) on a bounded sample of the same frame shots
    from oracle import c_oracle, philox
    ir = np.asarray(job.ir)
    noise = {int(k): (int(job.noise_channel[k]), philox.keep_threshold53(float(job.noise_prob[k])))
             for k in np.nonzero(ir[:, 0] == 17)[0]}
    pairs = [(int(r['measurement_value']), bool(r['deterministic'])) for r in ref]
    n_s, t = 64, 0.0
    while True:
        t0 = time.perf_counter()
        c_oracle.frame(job.n, job.d, ir, pairs, n_s, job.seed, job.stream, noise, shot0=1)
        t = time.perf_counter() - t0
        if t > 0.25 * cpu_seconds or n_s >= job.shots - 1:
            break
        n_s = min(job.shots - 1, n_s * 4)
    out["cpu_baseline"] = {"value": ops_per_shot * n_s / t, "unit": "gate-shots/s", "cores": 1, "kind": "port",
                           "sample": f"oracle_frame (C restatement of sdim simulate_frame) on {n_s} frame shots "
                                     f"of the repetition-code circuit, {t:.2f} s"}
    return out


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus above 1 is not supported "
                                   "(sharding across GPUs is designed, not built)"}))
        sys.exit(2)
    cpu_par = None
    if not args.no_cpu and args.workload != "frame":
        # before anything touches the GPU (the worker processes are spawned from this one)
        procs = min(16, len(os.sched_getaffinity(0)))
        cpu_par = cpu_baseline_parallel(make_workload(args.workload).jobs, args.cpu_seconds / 2, procs)
    from sdim_amd.engine import Engine

    if args.workload == "frame":
        eng = Engine()
        print(json.dumps({"frame": frame_sampler(eng, args.steps, args.warmup, args.cpu_seconds)}), flush=True)
        return
    wl = make_workload(args.workload)
    jobs = wl.jobs
    gate_shots = wl.gate_shots
    opts = {}
    if args.shots_per_item:
        opts["shots_per_item"] = args.shots_per_item
    if args.threads:
        opts["threads"] = args.threads
    tuned = None
    options_source = "defaults"
    base_opts = dict(opts)       # the other legs are other workloads: no options tuned for this one
    from sdim_amd import autotune
    if args.tune:
        tuned = autotune.tune(jobs, reps=max(1, min(args.steps, 3)), max_seconds=150.0)
        opts.update(tuned["options"])
        options_source = "autotuned in this run"
    elif not args.defaults and not args.shots_per_item and not args.threads:
        shipped = autotune.cached(jobs)
        if shipped:
            opts.update(shipped)
            tuned = {"options": shipped, "default_ms": None}
            options_source = "sdim_amd/tuning.json (autotuner result for this workload)"
    eng = Engine(**opts)

    t_submit = time.perf_counter()
    tk = eng.submit(jobs, "batched").wait()
    t_submit = time.perf_counter() - t_submit
    st = tk.stats()

    if args.warmup:
        tk.replay(args.warmup)
    t0 = time.perf_counter()
    dev_ms = tk.replay(args.steps)
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps
    value = gate_shots / (wall / args.steps)

    # per-class kernel timing (HIP events on the launch stream)
    classes = {}
    for which in ("hbm", "lds"):
        n_cls = st["items"] - st["lds_items"] if which == "hbm" else st["lds_items"]
        if n_cls == 0:
            continue
        ms = tk.replay(args.steps, which) / args.steps
        s2 = tk.stats()
        classes[which] = {"ms": ms, "bytes": s2["bytes_hbm"] if which == "hbm" else s2["bytes_lds"], "items": n_cls}

    def roof(cls, peak):
        if cls not in classes:
            return None
        c = classes[cls]
        ach = c["bytes"] / (c["ms"] * 1e-3) / 1e9
        traffic = None
        if args.workload == "uneven":
            # HBM bytes (FETCH_SIZE + WRITE_SIZE) of the HBM class, LDS transfers of the
            # LDS-resident classes, both from the rocprofv3 counter runs in the profile
            traffic = (profiled_traffic("batch_kernel_uneven_batch") if cls == "hbm"
                       else profiled_value("batch_kernel_uneven_batch", "lds_traffic_bytes_per_launch"))
        return {"bound": "hbm" if cls == "hbm" else "lds", "achieved": round(ach, 1), "peak": peak,
                "unit": "GB/s", "frac": round(ach / peak, 4), "traffic": traffic,
                "traffic_source": "profiles/batch_kernel_uneven_batch.md" if traffic else None,
                "kernel": ("sdim::batch_kernel<%d>" % (eng.get_option("threads") or eng.get_option("hbm_threads"))
                           if cls == "hbm" else "sdim::batch_kernel<NT, true> (LDS-resident classes)"),
                "items": c["items"], "algorithmic_bytes": c["bytes"], "kernel_ms": round(c["ms"], 3)}

    out = {
        "metric": METRIC, "value": value, "unit": "gate-shots/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8,u64", "data": "synthetic",
        "config": {"workload": wl.name, "description": wl.description, "circuits": len(jobs),
                   "gate_shots": gate_shots, "items": st["items"], "lds_items": st["lds_items"],
                   # the X/Z blocks of a circuit do not depend on the shot: this many shots share one
                   # X/Z evolution (each keeps its own phase vector); the per_shot_tableaux leg runs
                   # every shot on its own tableau
                   "xz_shared_by_shots": sum(int(j.shots) for j in jobs) / st["items"],
                   "threads": eng.get_option("threads"), "shots_per_item": eng.get_option("shots_per_item"),
                   "options": ({k: eng.get_option(k) for k in tuned["options"]} if tuned else "defaults"),
                   "options_source": options_source,
                   "tuned_from_ms": tuned["default_ms"] if tuned else None,
                   "other_legs_options": base_opts or "defaults"},
        "device_ms_per_step": dev_ms / args.steps,
        "first_submit_s": t_submit,
        "roofline": roof("hbm", HBM_PEAK_GBS) or roof("lds", LDS_PEAK_GBS),
        "roofline_lds": roof("lds", LDS_PEAK_GBS),
        "dominant_kernel": max(classes, key=lambda k: classes[k]["ms"]) if classes else None,
    }
    tk.release()

    if not args.no_plain:
        sub = jobs[::args.plain_every] if len(jobs) > 1 else jobs
        from sdim_amd.workloads import count_gate_shots
        gs = count_gate_shots(sub)
        tp = eng.submit(sub, "plain").wait()
        t0 = time.perf_counter()
        tp.replay(1)
        t_plain = time.perf_counter() - t0
        tp.release()
        tb = eng.submit(sub, "batched").wait()
        tb.replay(1)
        t0 = time.perf_counter()
        tb.replay(args.steps)
        t_b = (time.perf_counter() - t0) / args.steps
        tb.release()
        out["plain"] = {"value": gs / t_plain, "unit": "gate-shots/s", "circuits": len(sub),
                        "gate_shots": gs, "seconds": t_plain,
                        "sample": f"every {args.plain_every}-th circuit of the workload",
                        "batched_same_subset": gs / t_b}
        # batched over plain on the same circuits (the plain path runs a subset only)
        out["speedup_vs_plain"] = (gs / t_b) / (gs / t_plain)

    if not args.no_per_shot and args.workload == "uneven":
        # every shot on its own tableau (X/Z evolved once per shot, not once per circuit)
        e1 = Engine(shots_per_item=1, **{k: v for k, v in opts.items() if k != "shots_per_item"})
        t1 = e1.submit(jobs, "batched").wait()
        t1.replay(1)
        t0 = time.perf_counter()
        t1.replay(args.steps)
        w1 = (time.perf_counter() - t0) / args.steps
        st1 = t1.stats()
        t1.release()
        e1.close()
        out["per_shot_tableaux"] = {"value": gate_shots / w1, "unit": "gate-shots/s", "ms_per_step": 1e3 * w1,
                                    "items": st1["items"], "xz_shared_by_shots": 1,
                                    "config": "shots_per_item=1: one tableau (X, Z and phases) per shot"}

    leg_eng = Engine(**base_opts) if base_opts != opts else eng
    if not args.no_large and args.workload == "uneven":
        out["large_tableau"] = large_tableau(leg_eng, max(1, min(args.steps, 5)), 1)

    if not args.no_frame and args.workload == "uneven":
        out["frame"] = frame_sampler(leg_eng, max(1, min(args.steps, 10)), 1, args.cpu_seconds)

    if not args.no_published and args.workload == "uneven":
        out["published_workload"] = published_workload(leg_eng, max(1, min(args.steps, 5)),
                                                        0 if args.no_cpu else args.cpu_seconds)

    if not args.no_serve:
        out["serving"] = serving(leg_eng, args.serve_requests, args.serve_rate, args.serve_budget_gb)

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(jobs, args.cpu_seconds)
        out["cpu_baseline_parallel"] = cpu_par
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
