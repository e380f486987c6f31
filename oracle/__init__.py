"""CPU oracle for the Chirp-Z De-Doppler Transform (ZDT) path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the
Python host mirror, the Julia host module) imports, links or calls anything in
this package.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and there only as the checker
(or, for ``cpu_baseline``, as the timed CPU restatement of the reference).

Modules
-------
zdt_oracle
    Transforms: the FP64 exact-phase definition of the FDR (``fdr64``), an FP64
    restatement of the reference's Bluestein ZDT (``zdt64``), an FP32
    restatement that follows the reference's Float32 promotion order step by
    step (``zdt_ref32``), brute-force ``intfdr``/``intshift`` and the FP64
    ``fftfdr``/``fdshift``.
host_oracle
    Host-side arithmetic: ``calcNl``, ``growNr``, ``estimate_memory``,
    ``batchrates``, ``fdrstats``, ``fdrnormalize``, ``fdrdenormalize`` and
    ``findprotohits``.  (``driftrates`` needs no restatement: the GPU test
    checks the product's ranges against hand-derived exact values.)
synth
    Seeded synthetic spectrograms (Exponential(1) noise plus drifting tones).

How each function is pinned (golden vectors, closed forms, invariants) is
listed in the module headers; every function has at least one such check in
``tests/test_oracle.py``.
"""
