"""Runs circuits through the CPU emulation of the kernels and compares them with the C oracle.

    python tests/cpu_sim/run_sim.py --flavor asan --case random --seed 3 --count 24
    python tests/cpu_sim/run_sim.py --flavor asan --case uneven --first 0 --stride 64 --check 8

Cases:
  random   random circuits with every gate id, noise, M / M_X / RESET (d in {2,3,5,7,11}),
           batched (LDS-resident and, with --hbm, HBM-resident items) and plain
  uneven   circuits of the 4096-circuit uneven batch (n ~ U[32,512], d ~ U{2,3,5,7}, 10n
           gates + measure-all, 8 shots), with the device's launch plan (256 CUs, 8 shots per
           item); every --stride-th circuit from --first, the first --check of them compared
           with the oracle
Exit status 0 when every compared record and tableau matches; a sanitizer report ends the
process with the sanitizer's status.
"""
import argparse
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--flavor", default="plain")
    ap.add_argument("--case", default="random", choices=["random", "uneven", "selftest", "noalloc", "snapshots"])
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--count", type=int, default=12)
    ap.add_argument("--hbm", action="store_true", help="random case: also with every tableau in HBM")
    ap.add_argument("--first", type=int, default=0)
    ap.add_argument("--stride", type=int, default=64)
    ap.add_argument("--limit", type=int, default=0, help="uneven case: at most this many circuits")
    ap.add_argument("--check", type=int, default=4, help="uneven case: circuits compared with the oracle")
    ap.add_argument("--digests", action="store_true",
                    help="uneven case: compare every circuit with tests/golden/uneven_batch_digests.npz")
    ap.add_argument("--opt", action="append", default=[], help="engine option key=value")
    ap.add_argument("--which", type=int, default=0, help="selftest case")
    ap.add_argument("--expect-error", default="", help="random case: every submission must fail with this message")
    ap.add_argument("--same", type=float, default=0.0,
                    help="random case: share of two-qudit gates whose two qudits coincide")
    a = ap.parse_args()
    os.environ.setdefault("SDIM_SIM_CUS", "256" if a.case == "uneven" else "4")
    from tests.cpu_sim.sim_engine import SimEngine
    from parity_helpers import check_job, check_ticket, rand_job

    opts = {k: int(v) for k, v in (o.split("=", 1) for o in a.opt)}
    if a.case == "selftest":
        import ctypes
        from tests.cpu_sim import build_sim
        L = ctypes.CDLL(build_sim.lib_path(a.flavor))
        print(f"[sim] selftest {a.which}: {L.sdim_sim_selftest(a.which)}", flush=True)
        return
    t0 = time.time()
    if a.case == "noalloc":
        import ctypes
        from noalloc_helpers import check_noalloc_against_submit
        eng = SimEngine(a.flavor, **opts)
        rng = np.random.default_rng(a.seed)

        def alloc(nbytes):
            buf = (ctypes.c_uint8 * max(nbytes, 1))()
            alloc.keep.append(buf)
            return ctypes.addressof(buf)
        alloc.keep = []

        def to_host(ptr, nbytes):
            return np.ctypeslib.as_array((ctypes.c_uint8 * nbytes).from_address(ptr)).copy()

        def fill(ptr, value, nbytes):
            ctypes.memset(ptr, value, nbytes)

        check_noalloc_against_submit(eng, rng, alloc, to_host, fill)
        print(f"[sim] noalloc ok ({time.time() - t0:.1f} s)", flush=True)
        eng.close()
        return
    if a.case == "snapshots":
        from snapshot_helpers import check_snapshots
        rng = np.random.default_rng(a.seed)
        eng = SimEngine(a.flavor, **opts)
        check_snapshots(eng, rng, a.count)
        print(f"[sim] snapshots ok ({time.time() - t0:.1f} s)", flush=True)
        eng.close()
        return
    if a.case == "random":
        rng = np.random.default_rng(a.seed)
        for d in (2, 3, 5, 7, 11):
            jobs = [rand_job(rng, int(rng.integers(1, 14)), d, int(rng.integers(1, 70)), stream=i, same=a.same)
                    for i in range(a.count)]
            variants = [("batched", dict(opts)), ("plain", dict(opts))]
            if a.hbm:
                variants.append(("batched", dict(opts, allow_lds=0)))
            for mode, o in variants:
                eng = SimEngine(a.flavor, **o)
                tk = eng.submit(jobs, mode=mode)
                if a.expect_error:
                    for _ in range(2):      # the error sticks to the ticket
                        try:
                            tk.wait()
                        except Exception as e:  # noqa: BLE001
                            assert a.expect_error in str(e), str(e)
                        else:
                            raise AssertionError("the submission did not fail")
                    tk._released = True
                else:
                    tk.wait()
                    check_ticket(tk, jobs)
                    tk.release()
                eng.close()
                print(f"[sim] d={d} {mode} {o} ok ({time.time() - t0:.1f} s)", flush=True)
    else:
        from sdim_amd import workloads as W
        wl = W.uneven_batch()
        idx = list(range(a.first, len(wl.jobs), a.stride))
        if a.limit:
            idx = idx[:a.limit]
        jobs = [wl.jobs[i] for i in idx]
        eng = SimEngine(a.flavor, **dict({"shots_per_item": 8}, **opts))
        tk = eng.submit(jobs)
        if a.expect_error:
            for _ in range(2):      # the error sticks to the ticket
                try:
                    tk.wait()
                except Exception as e:  # noqa: BLE001
                    assert a.expect_error in str(e), str(e)
                else:
                    raise AssertionError("the submission did not fail")
            print(f"[sim] uneven: {len(jobs)} circuits failed as expected", flush=True)
            tk._released = True
            return
        tk.wait()
        st = tk.stats()
        print(f"[sim] uneven: {len(jobs)} circuits (n {min(j.n for j in jobs)}..{max(j.n for j in jobs)}), "
              f"{st['items']} items, {st['lds_items']} in LDS, {time.time() - t0:.1f} s", flush=True)
        if a.digests:
            from parity_helpers import digest_records, digest_tableau, records_array
            g = np.load(os.path.join(ROOT, "tests", "golden", "uneven_batch_digests.npz"))
            bad = [i for k, i in enumerate(idx)
                   if digest_records(records_array(tk.records(k))) != g["records"][i].tobytes()
                   or digest_tableau(tk.tableau(k, 0)) != g["tableau0"][i].tobytes()]
            print(f"[sim] digests: {len(idx) - len(bad)} of {len(idx)} circuits match", flush=True)
            if bad:
                raise SystemExit(f"circuits differing from the oracle: {bad[:16]}")
        for k in range(min(a.check, len(jobs))):
            check_job(tk, k, jobs[k], tableaux=True, shots=(0, jobs[k].shots - 1))
            print(f"[sim] circuit {idx[k]} (n={jobs[k].n}, d={jobs[k].d}) matches the oracle", flush=True)
        tk.release()
        eng.close()
    print(f"[sim] done in {time.time() - t0:.1f} s", flush=True)


if __name__ == "__main__":
    main()
