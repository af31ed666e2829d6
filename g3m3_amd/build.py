"""Builds the HIP library libg3m3.so in-tree for gfx950.

    python -m g3m3_amd.build          (or build() from __graft_entry__)

Every .hip file under g3m3_amd/csrc is compiled with hipcc and linked into
g3m3_amd/libg3m3.so.  Flags:
  -ffp-contract=off      keep each multiply and add separately rounded (the
                         kernels reproduce the reference's single-precision
                         expression order);
  -target-feature -packed-fp32-ops   no packed FP32 instructions.

When flang is present the Fortran driver fragments g3m3_amd/fortran/
drive_fields.f90 and drive_bounds.f90 (reference-style calls through
implicit interfaces) and drive_iface.f90 (the same calls through the bind(c)
interface module g3m3_iface.f90) are linked against the library beside
their sources.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "libg3m3.so"
OBJ = PKG / "_build"
ARCH = os.environ.get("G3M3_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
FLAGS = ["-O3", f"--offload-arch={ARCH}", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument", "-I", str(PKG.parent / "include")]


def _compile(src: Path) -> Path:
    obj = OBJ / (src.stem + ".o")
    cmd = [HIPCC, *FLAGS, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose: bool = True) -> Path:
    OBJ.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,--no-undefined", "-o",
           str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT} from {len(srcs)} sources")
    _build_fortran(verbose)
    return OUT


FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")


def _build_fortran(verbose: bool) -> None:
    iface = PKG / "fortran" / "g3m3_iface.f90"
    for name in ("drive_fields", "drive_bounds", "drive_iface"):
        src = PKG / "fortran" / (name + ".f90")
        exe = src.with_suffix("")
        if not (src.exists() and os.access(FLANG, os.X_OK)):
            continue
        srcs = [str(iface), str(src)] if name == "drive_iface" else [str(src)]
        cmd = [FLANG, "-O2", *srcs, "-L", str(PKG), "-lg3m3", "-Wl,-rpath,$ORIGIN/..",
               "-o", str(exe)]
        r = subprocess.run(cmd, capture_output=True, text=True, cwd=str(OBJ))
        if r.returncode != 0:
            raise RuntimeError(f"flang failed on {src.name}:\n{r.stderr[-4000:]}")
        if verbose:
            print(f"built {exe}")


if __name__ == "__main__":
    build()
    sys.exit(0)
