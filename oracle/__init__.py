"""CPU FP64 oracle for the gradient-domain Poisson (MG-PCG) path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (the ``gdpoisson`` package and
``libgdpoisson.so``) imports, links or calls this package.  It is used by:

* ``tests/`` as the checker the GPU results are compared against;
* ``__graft_entry__.smoke()`` as the checker of its one small GPU solve;
* ``bench.py``'s ``cpu_baseline`` leg, which times it on the host cores.

Contents
--------
``operators``   restatement of ``gradient-domain/src/gradient_domain/operators.py``.
``multigrid``   restatement of ``multigrid.py`` (Dirichlet, vertex-centred) plus the
                Neumann (cell-centred) V-cycle that the reference does not have.
``solvers``     restatement of ``solvers.py`` (registry, MG-PCG, DCT Neumann solve)
                plus the Neumann MG-PCG and an FP32-V-cycle / FP64-outer variant.
``generators``  the reference's manufactured problem (``benchmark.py:65-95``) and the
                seeded workload generators of the solve configurations.

Parity pins (see ``tests/test_oracle_*.py``): every function is checked against a
known answer that does not restate its own formula -- the reference's published
iteration counts and residuals (``gradient-domain/docs/benchmark.json``), the
reference tests' closed forms and thresholds (``gradient-domain/tests/``), dense
brute-force matrices, adjointness identities and the exact DCT solve.
"""
