"""CPU oracle for the prime-dimension stabilizer tableau engine.

TEST INFRASTRUCTURE ONLY.  Nothing in the product path (``sdim_amd``) imports,
links or executes anything under this directory.  Only the test-suite, the
``smoke()`` check in ``__graft_entry__`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as the checker / reported CPU baseline.

Contents
--------
philox.py    Philox4x64-10 counter-based generator (pinned against numpy's
             ``numpy.random.Philox``) and the draw keying shared with the
             device engine (measurement outcome, noise coin, noise exponents).
tableau.py   Literal restatement of sdim's ``ExtendedTableau`` and its flat
             gate sweeps (``sdim/tableau/tableau_prime.py``,
             ``sdim/tableau/tableau_optimized.py``), int64 numpy arrays,
             lazy reduction exactly as the reference does it.
gates.py     Gate-id dispatch with the reference's primitive words
             (``sdim/tableau/tableau_gates.py`` prime branches).
program.py   Shot loop of ``Program._simulate_tableau``
             (``sdim/program.py:269-365``) with Philox draws and tableau
             Pauli noise; returns records and final tableaux.
statevector.py  Brute-force dense statevector model built from the matrix
             formulas of ``sdim/unitary.py`` - used only to pin the oracle.
sdim_oracle.c   The same algorithm in plain C with eager mod-d reduction
             (fast checker for medium/large n and the CPU baseline "port").
"""
