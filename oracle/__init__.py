"""CPU oracle for the standalone SVD hot path -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.

It checks the GPU path and is never shipped as part of it.  Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` may import it, and only to check or to time the CPU restatement beside the
GPU path.  The product package never imports it and never falls back to it.

Contents
--------
``restatement``  numpy restatement of ``svdbench/standalone`` (``primitives.py`` and ``solvers.py``)
                 for the one-sided Jacobi, QDWH polar and TSQR families, operation by operation,
                 vectorised over the batch exactly as the reference is.
``metrics``      restatement of ``svdbench/metrics.py`` (FP64 LAPACK ``gesvd`` oracle and gates).
``matrices``     restatement of ``svdbench/matrices.py`` plus the seeded generators for the five
                 benchmark workloads.

Pinning (each function is checked by ``tests/test_oracle.py`` against something other than
itself): reference-held golden data (corpus inputs with the sweep / iteration / leaf counts the
reference recorded for them), SciPy/LAPACK (``geqrf`` R, ``potrf``, ``trtrs``, ``syevd``,
``gesvd``), closed forms (tournament schedule, QDWH weight equioscillation) and brute force on
tiny inputs.  Functions without such a check: none.
"""
