"""``Program``: the drop-in driver of sdim's tableau simulation on the GPU engine.

Mirrors ``sdim.program.Program`` (sdim/program.py:178-573): same constructor,
``simulate`` keyword arguments and ``SimulationOptions`` fields, same return
shapes (a flat list ordered by qudit then round for one shot, otherwise
``[qudit][round][shot]``), ``measurement_results`` instance state, and the
``_results_to_array`` / ``_combine_results`` / ``_build_ir`` helpers with the
reference's formats.

Dispatch as in the reference (program.py:244-267): with ``shots > 1`` and
neither ``record_tableau`` nor ``force_tableau``, one noiseless reference shot
runs the tableau algorithm (N1 acts as the identity there, program.py:31) and
the other ``shots - 1`` shots run the device Pauli-frame sampler
(program.py:45-165); otherwise every shot runs the tableau algorithm
(program.py:269-365) on the device.  Random outcomes, frame randomisation and
N1 noise are drawn from Philox4x64-10 keyed by (seed, stream) and (instruction
index, shot), so a run is reproducible from its seed.  Unlike the reference's
tableau path, the device tableau path applies N1 noise (as X^a Z^b) in every
shot it simulates.
"""
import copy
import random
from dataclasses import dataclass
from typing import Optional

import numpy as np

from .circuit import Circuit, ops_to_ir
from .engine import MEASUREMENT_DTYPE, Job, default_engine
from .tableau import ExtendedTableau, MeasurementResult, Tableau, is_prime


@dataclass
class SimulationOptions:
    shots: int = 1
    show_measurement: bool = False
    record_tableau: bool = False
    force_tableau: bool = False
    verbose: bool = False
    show_gate: bool = False
    exact: bool = False
    mode: str = "batched"      # "batched" (persistent kernel) or "plain" (one launch per op)
    seed: Optional[int] = None


class Program:
    def __init__(self, circuit: Circuit, tableau=None, engine=None, seed=None):
        if tableau is None:
            if not is_prime(circuit.dimension):
                raise NotImplementedError(
                    f"dimension {circuit.dimension} is composite; the MI355X engine covers prime dimensions only")
            self.stabilizer_tableau = ExtendedTableau(circuit.num_qudits, circuit.dimension)
        else:
            self.stabilizer_tableau = tableau
        self.circuits = [circuit]
        self.measurement_results = []
        self.initial_tableau = copy.copy(self.stabilizer_tableau)
        self.engine = engine
        self.seed = seed
        self.last_seed = None
        self.last_stream = 0

    # ------------------------------------------------------------ simulation
    def simulate(self, shots: int = 1, show_measurement: bool = False, record_tableau: bool = False,
                 force_tableau: bool = False, verbose: bool = False, show_gate: bool = False,
                 exact: bool = False, options: SimulationOptions = None):
        if options is None:
            options = SimulationOptions(shots=shots, show_measurement=show_measurement,
                                        record_tableau=record_tableau, force_tableau=force_tableau,
                                        verbose=verbose, show_gate=show_gate, exact=exact)
        if options.shots > 1 and not options.record_tableau and not options.force_tableau:
            return self._simulate_frame(options)
        return self._simulate_tableau(options)

    def _operations(self):
        ops = []
        for c in self.circuits:
            ops.extend(c.operations)
        return ops

    def _job(self, ops, shots, seed, stream=0, noiseless=False):
        tab = self.initial_tableau
        n, d = tab.num_qudits, tab.dimension
        ir, chan, prob = ops_to_ir(ops)
        if noiseless:                 # N1 -> I, keeping instruction indices (Philox counters)
            ir[ir[:, 0] == 17, 0] = 0
        init = None
        if not (isinstance(tab, ExtendedTableau) and tab.is_initial_state()):
            if not isinstance(tab, ExtendedTableau):
                raise NotImplementedError("an initial tableau must be an ExtendedTableau")
            init = tuple(np.asarray(a, dtype=np.int64) % m for a, m in
                         zip(tab.arrays(), (d, d, tab.order, d, d, tab.order)))
        has_noise = bool((ir[:, 0] == 17).any()) if len(ir) else False
        return Job(n, d, ir, shots=shots, seed=seed, stream=stream,
                   noise_channel=chan if has_noise else None, noise_prob=prob if has_noise else None,
                   init=init)

    def _simulate_tableau(self, options: SimulationOptions):
        engine = self.engine or default_engine()
        ops = self._operations()
        n = self.stabilizer_tableau.num_qudits
        d = self.stabilizer_tableau.dimension
        shots = int(options.shots)
        if shots < 1:
            raise ValueError("shots must be >= 1")
        seed = options.seed if options.seed is not None else (self.seed if self.seed is not None
                                                              else random.getrandbits(63))
        self.last_seed = seed
        job = self._job(ops, shots, seed, self.last_stream)
        # record_tableau: the kernel copies the tableau at every measurement event of this run
        tk = self._run_with_snapshots(engine, job, options.mode, 1 if options.record_tableau else 0)
        rec = tk.records(0)                       # (events, shots)
        final = tk.tableau(0, shots - 1)
        snapshots = None
        if options.record_tableau:
            snapshots = self._measurement_snapshots(tk, n, d, rec.shape[0], shots)
        tk.release()
        if options.show_gate or options.verbose:
            self._trace(engine, ops, seed, options)

        self.measurement_results = [[] for _ in range(n)]
        for ev in range(rec.shape[0]):
            q = int(rec[ev, 0]['qudit_index'])
            rnd = int(rec[ev, 0]['meas_round'])
            dets = rec[ev]['deterministic']
            vals = rec[ev]['measurement_value']
            while len(self.measurement_results[q]) <= rnd:
                self.measurement_results[q].append([])
            lst = self.measurement_results[q][rnd]
            for s in range(shots):
                mr = MeasurementResult(q, bool(dets[s]), int(vals[s]))
                if snapshots is not None:
                    mr.stabilizer_tableau = snapshots[ev][s]
                lst.append(mr)

        self.stabilizer_tableau = ExtendedTableau(n, d)
        self.stabilizer_tableau.set_arrays(final)
        if options.show_measurement:
            for s in range(shots):
                print(f"Measurement results for shot {s + 1}:")
                for per_q in self.measurement_results:
                    for rnd, shots_list in enumerate(per_q):
                        print(shots_list[s])

        if shots == 1:
            flat = []
            for per_q in self.measurement_results:
                for shots_list in per_q:
                    flat.append(shots_list[0])
            return flat
        return self.measurement_results

    def _simulate_frame(self, options: SimulationOptions):
        """sdim/program.py:244-265: a noiseless tableau reference shot (shot 0), then
        ``shots - 1`` Pauli-frame shots (shots 1..) on the device, combined per qudit and
        round as by ``_combine_results``."""
        engine = self.engine or default_engine()
        tab = self.initial_tableau
        if not (isinstance(tab, ExtendedTableau) and tab.is_initial_state()):
            raise NotImplementedError("the Pauli-frame sampler starts from |0...0>; use force_tableau=True "
                                      "with an initial tableau")
        ops = self._operations()
        n, d = tab.num_qudits, tab.dimension
        shots = int(options.shots)
        seed = options.seed if options.seed is not None else (self.seed if self.seed is not None
                                                              else random.getrandbits(63))
        self.last_seed = seed
        if options.show_gate or options.verbose:
            self._trace(engine, ops, seed, options, noiseless=True)
        # one submission: the noiseless reference shot (shot 0, on the tableau engine) and the
        # frame shots 1..shots-1 run concurrently; records come back combined
        tk = engine.frame_run([self._job(ops, shots, seed, self.last_stream)])
        rec = tk.records(0)                       # (events, shots), shot 0 = the reference shot
        final = tk.tableau(0, 0)
        tk.release()
        ref = rec[:, 0]
        rec = rec[:, 1:]

        self.measurement_results = [[] for _ in range(n)]
        for ev in range(ref.shape[0]):
            q = int(ref[ev]['qudit_index'])
            rnd = int(ref[ev]['meas_round'])
            while len(self.measurement_results[q]) <= rnd:
                self.measurement_results[q].append([])
            lst = self.measurement_results[q][rnd]
            lst.append(MeasurementResult(q, bool(ref[ev]['deterministic']), int(ref[ev]['measurement_value'])))
            dets = rec[ev]['deterministic']
            vals = rec[ev]['measurement_value']
            lst.extend(MeasurementResult(q, bool(dets[s]), int(vals[s])) for s in range(shots - 1))

        self.stabilizer_tableau = ExtendedTableau(n, d)
        self.stabilizer_tableau.set_arrays(final)
        if options.show_measurement:
            # the reference prints the reference shot only (program.py:351-353 with shots = 1)
            print("Measurement results for shot 1:")
            for per_q in self.measurement_results:
                for shots_list in per_q:
                    print(shots_list[0])
        return self.measurement_results

    @staticmethod
    def _run_with_snapshots(engine, job, mode, snapshots):
        """Runs one job with the engine's tableau snapshots set to `snapshots` for this
        submission (0 none, 1 at every measurement event, 2 after every instruction)."""
        if not snapshots:
            return engine.run([job], mode=mode)
        prev = engine.get_option("snapshots")
        engine.set_option("snapshots", snapshots)
        try:
            return engine.run([job], mode=mode)
        finally:
            engine.set_option("snapshots", prev)

    @staticmethod
    def _measurement_snapshots(tk, n, d, n_events, shots):
        """Tableau right after each measurement event (before RESET's X fix-up), per shot
        (sdim/program.py:322-324), as copied by the kernel at the event."""
        out = []
        for ev in range(n_events):
            per_shot = []
            for s in range(shots):
                t = ExtendedTableau(n, d)
                t.set_arrays(tk.snapshot(0, ev, s))
                per_shot.append(t)
            out.append(per_shot)
        return out

    def _trace(self, engine, ops, seed, options, noiseless=False):
        """``show_gate`` / ``verbose`` output of the first shot (sdim/program.py:313-350)."""
        n, d = self.stabilizer_tableau.num_qudits, self.stabilizer_tableau.dimension
        length = len(ops)
        if options.verbose:
            print("Initial state")
            self.initial_tableau.print_tableau()
            print("\n")
        tk = None
        if options.verbose:
            # one run of the first shot; the kernel copies the tableau after every instruction
            tk = self._run_with_snapshots(engine, self._job(ops, 1, seed, self.last_stream, noiseless),
                                          options.mode, 2)
        for k, op in enumerate(ops):
            if options.show_gate:
                info = op.target_index if op.target_index is not None else ""
                print("Time step" if k < length - 1 else "Final step", k, "\t", op.name, op.qudit_index, info)
            if options.verbose:
                t = ExtendedTableau(n, d)
                t.set_arrays(tk.snapshot(0, k, 0))
                t.print_tableau()
                print("\n")
        if tk is not None:
            tk.release()

    # ------------------------------------------------------------ formats (reference helpers)
    @staticmethod
    def _results_to_array(measurements: list) -> np.ndarray:
        """sdim/program.py:387-421."""
        if not measurements or not measurements[0]:
            raise ValueError("Empty or invalid measurement results format")
        first = measurements[0][0]
        max_rounds = max(len(m) for m in measurements)
        out = np.empty((len(measurements), max_rounds), dtype=MEASUREMENT_DTYPE)
        if isinstance(first, list):
            for q, per_q in enumerate(measurements):
                for r, shots_list in enumerate(per_q):
                    m = shots_list[0]
                    out[q, r] = (m.qudit_index, r, 0, m.deterministic, m.measurement_value)
            return out
        if isinstance(first, MeasurementResult):
            for q, per_q in enumerate(measurements):
                for r, m in enumerate(per_q):
                    out[q, r] = (m.qudit_index, r, 0, m.deterministic, m.measurement_value)
            return out
        raise ValueError("Invalid measurement results format")

    def _combine_results(self, frame_results) -> list:
        """sdim/program.py:423-454: append extra shots from a (qudit, round, shot) record array."""
        extra = frame_results.shape[2]
        for q in range(frame_results.shape[0]):
            for r in range(len(self.measurement_results[q])):
                self.measurement_results[q][r].extend(
                    MeasurementResult(int(frame_results[q, r, s]['qudit_index']),
                                      bool(frame_results[q, r, s]['deterministic']),
                                      int(frame_results[q, r, s]['measurement_value']))
                    for s in range(extra))
        return self.measurement_results

    def _build_ir(self, circuits, extra_shots: int):
        """sdim/program.py:456-526: IR triples without identity gates, and per-N1 noise samples
        (a, b) of shape (num_noise_gates, extra_shots, 2) drawn with numpy."""
        ir_list, noise_list = [], []
        d = self.stabilizer_tableau.dimension
        for circuit in circuits:
            for ins in circuit.operations:
                if ins.gate_id == 0:
                    continue
                t = ins.target_index if ins.target_index is not None else -1
                ir_list.append((ins.gate_id, ins.qudit_index, t))
                if ins.gate_id == 17:
                    params = ins.params or {}
                    channel = params.get("noise_channel", params.get("channel", "d"))
                    if channel == "d":
                        r = np.random.randint(1, d ** 2, size=extra_shots)
                        a, b = r % d, r // d
                    elif channel == "f":
                        a = np.random.randint(1, d, size=extra_shots)
                        b = np.zeros(extra_shots, dtype=np.int64)
                    elif channel == "p":
                        a = np.zeros(extra_shots, dtype=np.int64)
                        b = np.random.randint(1, d, size=extra_shots)
                    else:
                        raise ValueError("Must specify a valid noise channel.")
                    keep = np.random.uniform(0.0, 1.0, size=extra_shots) < 1.0 - float(params.get("prob", 0.01))
                    a = np.where(keep, 0, a)
                    b = np.where(keep, 0, b)
                    noise_list.append(np.stack((a, b), axis=1))
        ir_dtype = np.dtype([('gate_id', np.int64), ('qudit_index', np.int64), ('target_index', np.int64)])
        ir = np.array(ir_list, dtype=ir_dtype)
        noise = np.array(noise_list, dtype=np.int64) if noise_list else np.empty((1, extra_shots, 2), dtype=np.int64)
        return ir, noise

    # ------------------------------------------------------------ program management
    def append_circuit(self, circuit: Circuit):
        """sdim/program.py:528-544."""
        if self.circuits[-1].num_qudits < circuit.num_qudits:
            self.circuits[-1].num_qudits = circuit.num_qudits
        else:
            circuit.num_qudits = self.circuits[-1].num_qudits
        if self.circuits[-1].dimension != circuit.dimension:
            raise ValueError("Circuits must have the same dimension")
        self.circuits.append(circuit)

    def print_measurements(self):
        """Print the stored results (the reference re-simulates for a single shot,
        sdim/program.py:561-563; here the stored results are printed)."""
        shot_count = 0
        for per_q in self.measurement_results:
            for group in per_q:
                shot_count = max(shot_count, len(group))
        if shot_count == 0:
            print("No measurements recorded.")
            return
        if shot_count == 1:
            for per_q in self.measurement_results:
                for group in per_q:
                    print(group[0])
            return
        for s in range(shot_count):
            print(f"Shot {s + 1}:")
            for per_q in self.measurement_results:
                for r, group in enumerate(per_q):
                    print(f"{group[s]} during measurement {r}")
            print()

    def __str__(self) -> str:
        return str(self.stabilizer_tableau)
