"""CPU oracle for the TT-ALS path (TEST INFRASTRUCTURE ONLY).

A from-scratch numpy/scipy restatement of the reference algorithm
(scikit_tt/solvers/sle.py, scikit_tt/solvers/evp.py,
scikit_tt/tensor_train.py, scikit_tt/utils.py).  Every function cites the
reference lines it follows.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker / the timed CPU
baseline.  The product path (``scikit_tt_amd``) never imports it and has no
CPU fallback.

Parity pins (which check fixes each function) are listed in
``tests/test_oracle.py``; every function here is pinned by at least one of:
golden vectors written by the reference itself (``tests/golden``), a dense
brute-force computation, or a closed form.
"""
