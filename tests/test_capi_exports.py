"""The C-ABI library loads and exports every entry point include/g3m3.h
declares (no GPU needed: nothing is called)."""
import ctypes
import os

import pytest

from g3m3_amd import _lib


def test_header_symbols_exported():
    if not _lib.LIB_PATH.exists():
        pytest.skip("libg3m3.so not built")
    so = ctypes.CDLL(str(_lib.LIB_PATH), mode=os.RTLD_LOCAL)
    names = _lib.exported_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_header_matches_reference_signatures():
    """Every entry point named after a reference subroutine takes that
    subroutine's dummies in order (checked by count against the oracle's
    argument lists, which mirror the reference declarations)."""
    import re
    from oracle import argnames as A
    hdr = (_lib.LIB_PATH.parent.parent / "include" / "g3m3.h").read_text()
    decl = {m.group(1): m.group(2) for m in
            re.finditer(r"\b(g3m3_[a-z0-9_]+)\s*\(([^;]*)\);", hdr, re.S)}

    def nargs(name):
        body = decl[name].strip()
        return 0 if body in ("", "void") else body.count(",") + 1
    assert nargs("g3m3_push_ion") == len(A.push_ion_names("", "old", "wrk", "q", "rmassq"))
    assert nargs("g3m3_set_rho") == len(A.set_rho_names(""))
    assert nargs("g3m3_set_speed_agrd") == len(A.set_speed_names(""))
    assert nargs("g3m3_bndry_outer") == len(A.bndry_outer_names(""))
    assert nargs("g3m3_bndry_inner") == len(A.bndry_inner_names("")) + 2   # + v_rot, r_rot
    assert nargs("g3m3_flux_correct") == len(A.flux_correct_names())
    assert nargs("g3m3_lap_plasma") == len(A.lap_plasma_names("q", "", "wrk"))


def test_fortran_names_exported():
    """Each subroutine entry point is also exported under the name an
    implicit-interface Fortran call uses (lower case + '_'), as an alias of
    the same code."""
    if not _lib.LIB_PATH.exists():
        pytest.skip("libg3m3.so not built")
    so = ctypes.CDLL(str(_lib.LIB_PATH), mode=os.RTLD_LOCAL)
    for n in ("g3m3_push_ion", "g3m3_calcur", "g3m3_bande", "g3m3_flux_correct",
              "g3m3_bndry_flanks_v", "g3m3_divb_correct", "g3m3_trilin_interp"):
        a = ctypes.cast(getattr(so, n), ctypes.c_void_p).value
        b = ctypes.cast(getattr(so, n + "_"), ctypes.c_void_p).value
        assert a == b, n
    for n in ("g3m3_init_", "g3m3_register_", "g3m3_to_device_", "g3m3_to_host_"):
        assert hasattr(so, n), n
