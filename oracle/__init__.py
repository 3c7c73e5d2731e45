"""CPU oracle for the skeletonized FMM apply -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package, and only as the checker (or, for the
baseline, as the timed CPU restatement). The product package ``skelfmm_amd``
never imports it.

Contents (each function cites the reference file:line it restates):

* ``kernels``        closed forms of the eight operators (``skelfmm/util.py:128-377``,
                     torch clamps of ``skelfmm/util_batched.py:141-185,240-656``) and a
                     chunked FP64 direct sum;
* ``serial_apply``   the per-pair serial matvec (``skelfmm/recursive_skel.py:784-1084``);
* ``batched_apply``  the padded, chunked torch-CPU matvec (``skelfmm/batched_fmm.py:92-564``),
                     the CPU baseline;
* ``id``             the pivoted-QR interpolative decomposition (``skelfmm/util.py:13-51``,
                     ``87-124``) in the reference's stacked form;
* ``p2p_ref.c``      a C restatement of the dense kernel sums for large direct sums.

Pinning (see the tests): every kernel against the reference tests' known
answers, adjoint identities and coincident-point rules; the ID against
reconstruction and eps/fixed-rank consistency; the two apply restatements
against each other (1e-12) and against the direct sum (below tol).
"""
