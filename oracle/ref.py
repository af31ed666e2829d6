"""ctypes access to the reference's own Fortran subroutines (oracle/_ref).

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``g3m3_amd``) imports
this module; it is used by ``tests/``, by ``__graft_entry__.smoke()`` and by
``bench.py``'s ``cpu_baseline`` leg, and only as a checker or as the timed
CPU baseline.

``oracle/Makefile`` compiles the reference's ``code/sub_*.f90`` files,
unchanged, with flang into ``oracle/_ref/libg3m3_ref.so`` (serial) and
``libg3m3_ref_omp.so`` (OpenMP).  The subroutines have implicit interfaces
(every argument passed by reference), so each call here passes:

* numpy arrays (C-contiguous, float32 or int32) as their first-element
  address -- a (ngrd, nz, ny, nx) C-order array is byte-identical to the
  reference's Fortran ``(nx, ny, nz, ngrd)``;
* Python ints as ``int32`` by reference, Python/numpy floats as ``real*4``
  by reference, ``F64(x)`` as ``real*8`` by reference, bools as flang
  ``LOGICAL*4`` (1 = .true.).

The common block ``/rotation/`` (code/3d_main.f90:633-634; read by
code/sub_bndry_inner.f90:62 and code/sub_dipole.f90:6) is reached through
the ``rotation_`` symbol of the shared object.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
REF_DIR = _HERE / "_ref"


class F64:
    """Marks a scalar that the Fortran side declares ``real(dp)``."""

    def __init__(self, v):
        self.v = float(v)


class Chars:
    """A ``CHARACTER*n`` actual argument: its bytes by reference, and its
    length appended after the last argument (the hidden-length convention
    of flang and gfortran)."""

    def __init__(self, text: str, n: int):
        self.b = ctypes.create_string_buffer(text.ljust(n)[:n].encode(), n)
        self.n = n


class _Rotation(ctypes.Structure):
    # common /rotation/v_rot,r_rot,rot_angle,xdip,ydip,zdip,sin_tilt,cos_tilt,b0
    _fields_ = [(n, ctypes.c_float) for n in
                ("v_rot", "r_rot", "rot_angle", "xdip", "ydip", "zdip",
                 "sin_tilt", "cos_tilt", "b0")]


def available(omp: bool = False) -> bool:
    return (REF_DIR / ("libg3m3_ref_omp.so" if omp else "libg3m3_ref.so")).exists()


class RefLib:
    """One loaded copy of the reference build."""

    def __init__(self, omp: bool = False):
        name = "libg3m3_ref_omp.so" if omp else "libg3m3_ref.so"
        path = REF_DIR / name
        if not path.exists():
            raise FileNotFoundError(
                f"{path} is not built; run `make -C oracle` where the "
                "reference sources are present")
        # RTLD_LOCAL + a private handle: two builds can coexist in-process
        self.lib = ctypes.CDLL(str(path), mode=os.RTLD_LOCAL)
        self.omp = omp
        self._keep = []

    # -- argument marshalling -------------------------------------------
    def _arg(self, a):
        if isinstance(a, np.ndarray):
            if not a.flags["C_CONTIGUOUS"]:
                raise ValueError("reference arrays must be C-contiguous")
            if a.dtype not in (np.float32, np.int32, np.float64):
                raise TypeError(f"unsupported dtype {a.dtype}")
            return ctypes.c_void_p(a.ctypes.data)
        if isinstance(a, F64):
            c = ctypes.c_double(a.v)
        elif isinstance(a, (bool, np.bool_)):
            c = ctypes.c_int32(1 if a else 0)
        elif isinstance(a, (int, np.integer)):
            c = ctypes.c_int32(int(a))
        elif isinstance(a, (float, np.floating)):
            c = ctypes.c_float(np.float32(a))
        else:
            raise TypeError(f"cannot pass {type(a)} to Fortran")
        self._keep.append(c)
        return ctypes.byref(c)

    def call(self, name: str, *args):
        fn = getattr(self.lib, name.lower() + "_")
        fn.restype = None
        cargs, hidden = [], []
        for a in args:
            if isinstance(a, Chars):
                cargs.append(ctypes.cast(a.b, ctypes.c_void_p))
                hidden.append(ctypes.c_size_t(a.n))
            else:
                cargs.append(self._arg(a))
        cargs += hidden
        try:
            fn(*cargs)
        finally:
            self._keep.clear()

    def rotation(self) -> _Rotation:
        return _Rotation.in_dll(self.lib, "rotation_")

    def set_threads(self, n: int):
        """omp_set_num_threads for the OpenMP build."""
        if self.omp:
            f = self.lib.omp_set_num_threads if hasattr(self.lib, "omp_set_num_threads") \
                else ctypes.CDLL("libomp.so").omp_set_num_threads
            f(ctypes.c_int(n))


def with_big_stack(fn, *args, stack_mb: int = 1024, **kw):
    """Run ``fn`` on a thread with a large stack.

    The reference keeps full-box automatic arrays on the stack (13 of them
    in code/sub_set_speed_agrd.f90:60-61, 9 in code/sub_fcsmooth.f90:207-208,
    plus the 7-band matrix of code/sub_divB_correct.f90:47-49); at
    121x121x61 they overflow the default 8 MB main-thread stack."""
    import threading
    out, err = [], []

    def body():
        try:
            out.append(fn(*args, **kw))
        except BaseException as e:  # re-raised on the caller's thread
            err.append(e)

    old = threading.stack_size()
    threading.stack_size(stack_mb << 20)
    try:
        t = threading.Thread(target=body)
        t.start()
        t.join()
    finally:
        threading.stack_size(old)
    if err:
        raise err[0]
    return out[0] if out else None
