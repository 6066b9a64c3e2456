"""CPU oracle for the projector / VCD hot path -- TEST INFRASTRUCTURE ONLY.

This package restates, in plain numpy/scipy (FP64 by default), the algorithm
of the reference's hot path so that the HIP implementation in
``mbirtorch_amd`` can be checked against it.  Each function cites the
reference file:line it follows (paths relative to the reference checkout).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import anything from here, and only as the checker or the
timed CPU baseline.  The product path (``mbirtorch_amd``) never imports it
and fails loudly when its HIP library is missing.

Pinning (what fixes each restated function independently of the restatement
itself) is listed in each module's header.  A function with no independent
check says "parity unpinned" there.
"""
