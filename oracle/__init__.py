"""CPU FP64 oracle for the recycled-subspace sigma_min(C - zI) solver.

TEST INFRASTRUCTURE ONLY.  Nothing in the product path (``rls_amd``) imports,
links or executes anything under ``oracle/``; only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use it,
and only as the checker / the timed CPU restatement.

Modules
-------
``rls_oracle``  numpy/scipy restatement of the reference algorithm (psTriRecy,
                lobpcgSvd, RecycledSpace, fastSvdRaylRitz, recursion grids),
                function by function, each citing the reference file:line.
``truth``       ground-truth sigma_min by independent FP64 methods (dense SVD,
                inverse iteration on the Schur factor, sparse LU shift-invert).
"""
