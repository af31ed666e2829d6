"""The oracle: test infrastructure for G3M^3's per-timestep hot path.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg import this package, and only as the checker (or the
timed CPU baseline).  The product package ``g3m3_amd`` never imports it,
and its GPU path fails loudly instead of falling back to anything here.

* ``oracle.restate`` -- NumPy restatement of each reference subroutine,
  same names and argument lists, single precision in the reference's
  operation order.
* ``oracle.ref``     -- ctypes access to the reference's own Fortran,
  compiled unchanged by ``oracle/Makefile`` into ``oracle/_ref/``.
* ``oracle.boxstep`` -- the reference's box-step / full-step call sequence
  (code/3d_main.f90:1925-3095), driving either of the two above.
"""
