"""ctypes binding of libg3m3.so (the C-ABI declared in include/g3m3.h).

There is no fallback: if the library is missing or no GPU is visible, the
first call raises.  Argument marshalling follows the C-ABI's by-reference
convention:

* torch tensors (device memory) pass their data pointer;
* numpy arrays pass their host address -- used only for the small arrays
  the library reads on the host (grid extents, xspac, list counts, times,
  band positions);
* Python ints -> int32 by reference, floats -> float by reference,
  ``F64(x)`` -> double by reference, bools -> int32 (1 = .true.).
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libg3m3.so"


class G3M3Error(RuntimeError):
    pass


class F64:
    """A scalar the reference declares real(dp)."""

    def __init__(self, v):
        self.v = float(v)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise G3M3Error(f"{LIB_PATH} is not built (run `python -m g3m3_amd.build`)")
        _lib = ctypes.CDLL(str(LIB_PATH), mode=os.RTLD_GLOBAL)
        _lib.g3m3_last_error_msg.restype = ctypes.c_char_p
        _lib.g3m3_divb_correct_ex.restype = ctypes.c_int
        _lib.g3m3_init.restype = ctypes.c_int
        if _lib.g3m3_init() != 0:
            raise G3M3Error("g3m3_init failed: " + _lib.g3m3_last_error_msg().decode())
    return _lib


def _is_tensor(a):
    return type(a).__module__.startswith("torch") and hasattr(a, "data_ptr")


class ByValue:
    """A ctypes scalar passed by value (the C-ABI's few value arguments, such
    as the hidden CHARACTER length of g3m3_findgrid)."""

    def __init__(self, c):
        self.c = c


def marshal(a, keep):
    if isinstance(a, ByValue):
        return a.c
    if isinstance(a, bytes):    # a C string (file paths)
        keep.append(a)
        return ctypes.c_char_p(a)
    if _is_tensor(a):
        if not a.is_cuda:
            raise G3M3Error("device arrays must be CUDA/HIP tensors")
        if not a.is_contiguous():
            raise G3M3Error("device arrays must be contiguous")
        return ctypes.c_void_p(a.data_ptr())
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise G3M3Error("host arrays must be C-contiguous")
        keep.append(a)
        return ctypes.c_void_p(a.ctypes.data)
    if isinstance(a, F64):
        c = ctypes.c_double(a.v)
    elif isinstance(a, (bool, np.bool_)):
        c = ctypes.c_int32(1 if a else 0)
    elif isinstance(a, (int, np.integer)):
        c = ctypes.c_int32(int(a))
    elif isinstance(a, (float, np.floating)):
        c = ctypes.c_float(np.float32(a))
    elif isinstance(a, ctypes._SimpleCData) or isinstance(a, ctypes.Array):
        keep.append(a)
        return ctypes.byref(a)
    elif a is None:
        return ctypes.c_void_p(0)
    else:
        raise G3M3Error(f"cannot pass {type(a)} through the C-ABI")
    keep.append(c)
    return ctypes.byref(c)


def call(name: str, *args):
    """Call ``g3m3_<name>`` with reference-style arguments."""
    fn = getattr(lib(), "g3m3_" + name.lower())
    keep = []
    cargs = [marshal(a, keep) for a in args]
    r = fn(*cargs)
    return r


def check(where: str = ""):
    """Raise if the library latched an error since the last check."""
    L = lib()
    e = L.g3m3_last_error()
    if e:
        raise G3M3Error(f"{where}: g3m3 error {e}: {L.g3m3_last_error_msg().decode()}")


def sync():
    lib().g3m3_sync()
    check("sync")


def exported_symbols():
    """Names declared in include/g3m3.h (for the load/export check)."""
    import re
    hdr = Path(__file__).resolve().parent.parent / "include" / "g3m3.h"
    return sorted(set(re.findall(r"\b(g3m3_[a-z0-9_]+)\s*\(", hdr.read_text())))
