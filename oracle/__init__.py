"""CPU oracle for the PDHG reconstruction hot path (TEST INFRASTRUCTURE ONLY).

This package restates, in float64 NumPy, the arithmetic that the reference
reconstruction path runs through third-party libraries:

* ``oracle.nufft``   -- SigPy 0.1.27 ``nufft`` / ``nufft_adjoint`` (Kaiser-Bessel
  width 4, oversampling 1.25) as called at src/toporecon/nufft_backend.py:104-107
  and :112-116.
* ``oracle.pdhg``    -- the TVM / TVME / TVMW PDHG iterations of
  src/toporecon/algorithms/{tvm,tvme,tvmw}.py on one process (all MPI peers
  PROC_NULL, group size 1).
* ``oracle.wavelet`` -- the db1 echo wavelet of tvmw.py:312-375 and the
  ptwt ``wavedec3`` / ``waverec3`` db6 zero-mode transform used at
  tvmw.py:196-206.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker or the CPU baseline.  The
product path (``toporecon_hip``) never imports this package and raises when its
HIP library is missing.

Pinning (what fixes each function independently of its own formula) is listed
in each module's docstring.  A function whose docstring says "parity unpinned"
has no such independent check.
"""
