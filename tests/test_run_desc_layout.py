"""The Python mirror of g3m3_run_desc has the C struct's size and field
offsets (compiled with the host C compiler from include/g3m3.h)."""
import ctypes
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIELDS = ("nx", "state", "bx0", "numsrf", "grd_xmin", "xspac", "rmassq", "eta", "isotropic",
          "tilting", "tilt", "dtilt", "ut", "t_equiv", "planet_per", "xdip", "dip_b0", "r_inner")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no host C compiler")
def test_run_desc_layout(tmp_path):
    from g3m3_amd.driver import RunDesc
    src = tmp_path / "layout.c"
    body = "".join(f'    printf("{f} %zu\\n", offsetof(g3m3_run_desc, {f}));\n' for f in FIELDS)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "g3m3.h"\n'
                   'int main(void) {\n'
                   '    printf("size %zu\\n", sizeof(g3m3_run_desc));\n' + body +
                   '    return 0;\n}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True, timeout=120)
    out = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True,
                                                        text=True, check=True).stdout.split("\n")
               if line)
    assert int(out["size"]) == ctypes.sizeof(RunDesc)
    for f in FIELDS:
        assert int(out[f]) == getattr(RunDesc, f).offset, f


FM_FIELDS = ("nx", "nz", "rho", "mom", "epres", "b0", "rx", "delt", "reynolds", "diferg")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no host C compiler")
def test_field_model_desc_layout(tmp_path):
    from g3m3_amd.adjoint import FieldModelDesc
    src = tmp_path / "fm.c"
    body = "".join(f'    printf("{f} %zu\\n", offsetof(g3m3_field_model_desc, {f}));\n'
                   for f in FM_FIELDS)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "g3m3.h"\n'
                   'int main(void) {\n'
                   '    printf("size %zu\\n", sizeof(g3m3_field_model_desc));\n' + body +
                   '    return 0;\n}\n')
    exe = tmp_path / "fm"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True, timeout=120)
    out = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True,
                                                        text=True, check=True).stdout.split("\n")
               if line)
    assert int(out["size"]) == ctypes.sizeof(FieldModelDesc)
    for f in FM_FIELDS:
        assert int(out[f]) == getattr(FieldModelDesc, f).offset, f
