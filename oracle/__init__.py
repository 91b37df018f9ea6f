"""CPU oracle for the SU(2) DMRG hot path -- TEST INFRASTRUCTURE ONLY.

This package is a plain numpy restatement of the reference algorithm
(lingwangqs/SU2DMRG).  It exists to check the HIP product path: only the
tests, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker or the timed CPU
baseline.  The product library (``su2dmrg_amd``) never calls into it and
fails loudly when its HIP extension is missing.

Every function cites the reference file:line it restates.  How each part is
pinned (closed forms, sympy, brute force, exact diagonalisation) is listed in
the module headers; a function without such a check says "parity unpinned".
"""
