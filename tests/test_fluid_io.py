"""Fluid (restart) files: the library's reader and writer against the
Fortran runtime's own unformatted sequential records, written and read by
a Fortran program with the reference's record sequence
(code/3d_main.f90:972-1011, 3310-3349).  Host code only: no GPU."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from g3m3_amd import _lib

FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
HERE = os.path.dirname(os.path.abspath(__file__))


def _lib_nodev():
    if not _lib.LIB_PATH.exists():
        pytest.skip("libg3m3.so not built")
    return ctypes.CDLL(str(_lib.LIB_PATH), mode=os.RTLD_LOCAL)


@pytest.fixture(scope="module")
def fortran_rw(tmp_path_factory):
    if not os.access(FLANG, os.X_OK):
        pytest.skip("flang not available")
    d = tmp_path_factory.mktemp("frw")
    exe = d / "fluid_rw"
    r = subprocess.run([FLANG, "-O1", os.path.join(HERE, "fortran", "fluid_rw.f90"), "-o",
                        str(exe)], capture_output=True, text=True, cwd=str(d))
    assert r.returncode == 0, r.stderr
    return str(exe)


def _case(seed=3):
    rng = np.random.default_rng(seed)
    nx, ny, nz, ng, mb, msrf, mmid, mzero = 7, 6, 5, 2, 1, 9, 4, 3
    t = float(rng.uniform(0, 100))
    f = rng.standard_normal((37, ng, nz, ny, nx)).astype(np.float32)
    lists = dict(
        parm_srf=rng.standard_normal((msrf, 7, mb)).astype(np.float32),
        parm_mid=rng.standard_normal((mmid, 7, mb)).astype(np.float32),
        parm_zero=rng.standard_normal((mzero, 7, mb)).astype(np.float32),
        ijzero=rng.integers(1, 5, (mzero, 3, mb)).astype(np.int32),
        numzero=np.array([mzero - 1], np.int32),
        ijmid=rng.integers(1, 5, (mmid, 3, mb)).astype(np.int32),
        nummid=np.array([mmid], np.int32),
        ijsrf=rng.integers(1, 5, (msrf, 3, mb)).astype(np.int32),
        numsrf=np.array([msrf - 2], np.int32))
    dims = (nx, ny, nz, ng, mb, msrf, mmid, mzero)
    return dims, t, f, lists


BLOB_ORDER = ["parm_srf", "parm_mid", "parm_zero", "ijzero", "numzero", "ijmid", "nummid",
              "ijsrf", "numsrf"]


def _blob(t, f, lists):
    return (np.array([t], np.float64).tobytes() + f.tobytes()
            + b"".join(lists[k].tobytes() for k in BLOB_ORDER))


def _c_args(dims, t, f, lists):
    nx, ny, nz, ng, mb, msrf, mmid, mzero = dims
    I = lambda v: ctypes.byref(ctypes.c_int(v))
    ptrs = (ctypes.c_void_p * 37)(*[f[a].ctypes.data for a in range(37)])
    P = lambda a: ctypes.c_void_p(a.ctypes.data)
    return ([I(nx), I(ny), I(nz), I(ng), ptrs, I(mb), I(msrf), I(mmid), I(mzero),
             P(lists["parm_srf"]), P(lists["parm_mid"]), P(lists["parm_zero"]),
             P(lists["ijsrf"]), P(lists["ijmid"]), P(lists["ijzero"]),
             P(lists["numsrf"]), P(lists["nummid"]), P(lists["numzero"])], ptrs)


def test_write_matches_fortran_and_reads_back(fortran_rw, tmp_path):
    L = _lib_nodev()
    dims, t, f, lists = _case()
    spec = tmp_path / "spec"
    spec.write_text(" ".join(map(str, dims)) + "\n")
    # library writer
    args, keep = _c_args(dims, t, f, lists)
    tt = ctypes.c_double(t)
    rc = L.g3m3_fluid_write(str(tmp_path / "c.fluid").encode(), ctypes.byref(tt), *args)
    assert rc == 0
    # Fortran writer from the same data
    (tmp_path / "in.blob").write_bytes(_blob(t, f, lists))
    r = subprocess.run([fortran_rw, "w", str(spec), str(tmp_path / "in.blob"),
                        str(tmp_path / "f.fluid")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    # byte-identical files
    assert (tmp_path / "c.fluid").read_bytes() == (tmp_path / "f.fluid").read_bytes()
    # Fortran reads the library's file back to the same data
    r = subprocess.run([fortran_rw, "r", str(spec), str(tmp_path / "c.fluid"),
                        str(tmp_path / "out.blob")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert (tmp_path / "out.blob").read_bytes() == _blob(t, f, lists)
    # the library reads the Fortran file back to the same data
    f2 = np.zeros_like(f)
    l2 = {k: np.zeros_like(v) for k, v in lists.items()}
    args2, keep2 = _c_args(dims, 0.0, f2, l2)
    t2 = ctypes.c_double(0)
    rc = L.g3m3_fluid_read(str(tmp_path / "f.fluid").encode(), ctypes.byref(t2), *args2)
    assert rc == 0 and t2.value == t
    assert np.array_equal(f2, f)
    for k in lists:
        assert np.array_equal(l2[k], lists[k]), k


def test_read_rejects_wrong_sizes(tmp_path):
    L = _lib_nodev()
    dims, t, f, lists = _case(5)
    args, keep = _c_args(dims, t, f, lists)
    tt = ctypes.c_double(t)
    assert L.g3m3_fluid_write(str(tmp_path / "a.fluid").encode(), ctypes.byref(tt), *args) == 0
    nx, ny, nz, ng, mb, msrf, mmid, mzero = dims
    bad = (nx, ny, nz, ng, mb, msrf + 1, mmid, mzero)    # list sizes that do not match
    l2 = {k: np.zeros_like(v) for k, v in lists.items()}
    l2["parm_srf"] = np.zeros((msrf + 1, 7, mb), np.float32)
    l2["ijsrf"] = np.zeros((msrf + 1, 3, mb), np.int32)
    f2 = np.zeros_like(f)
    args2, keep2 = _c_args(bad, 0.0, f2, l2)
    t2 = ctypes.c_double(0)
    assert L.g3m3_fluid_read(str(tmp_path / "a.fluid").encode(), ctypes.byref(t2), *args2) == 2
    assert L.g3m3_fluid_read(str(tmp_path / "missing").encode(), ctypes.byref(t2), *args2) == 1
