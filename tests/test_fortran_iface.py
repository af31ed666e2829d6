"""The bind(c) interface module g3m3_amd/fortran/g3m3_iface.f90 (written by
make_iface.py from include/g3m3.h, with each dummy's rank taken from the
reference subroutine of the same name).

* Every subroutine entry point of the header that the module is meant to
  bind is in it, with the header's argument names in the header's order.
* A caller that passes the reference's arguments compiles; one that passes a
  scalar where the reference subroutine takes an array does not (the check
  the module exists for).  Compile-only, with flang."""
import os
import re
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FORT = os.path.join(ROOT, "g3m3_amd", "fortran")
IFACE = os.path.join(FORT, "g3m3_iface.f90")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
sys.path.insert(0, FORT)


def test_module_matches_header():
    import make_iface as M
    text = open(IFACE).read()
    bound = 0
    for ret, name, params in M.header_decls():
        if name in ("init", "last_error", "sync", "register", "unregister", "to_device",
                    "to_host") or name.startswith(M.SKIP_PREFIX):
            continue
        m = re.search(rf"(subroutine|function) g3m3_{name}\((.*?)\) bind\(c, name='g3m3_{name}'\)",
                      text, re.S)
        assert m, f"g3m3_{name} missing from the module"
        args = [a.strip() for a in m.group(2).replace("&", " ").split(",") if a.strip()]
        assert args == [p[0] for p in params], name
        bound += 1
    assert bound >= 40


CALLER = """program c
    use g3m3_iface
    implicit none
    integer :: nx = 5, ny = 5, nz = 5, n_grids = 1, box = 1
    real :: rx = 1.0
    real :: bx(5,5,5,1), by(5,5,5,1), bz(5,5,5,1), cx(5,5,5), cy(5,5,5), cz(5,5,5)
    real :: scalar_b
    call g3m3_calcur(%s, by, bz, nx, ny, nz, n_grids, box, cx, cy, cz, rx, rx, rx)
end program
"""


@pytest.mark.skipif(not os.access(FLANG, os.X_OK), reason="no flang")
@pytest.mark.parametrize("arg,ok", [("bx", True), ("scalar_b", False)])
def test_module_checks_ranks(tmp_path, arg, ok):
    src = tmp_path / "c.f90"
    src.write_text(CALLER % arg)
    r = subprocess.run([FLANG, "-fsyntax-only", IFACE, str(src)], capture_output=True,
                       text=True, cwd=str(tmp_path), timeout=300)
    assert (r.returncode == 0) == ok, r.stderr[-2000:]
