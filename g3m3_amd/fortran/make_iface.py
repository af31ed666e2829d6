"""Writes g3m3_iface.f90: a Fortran module of bind(c) interfaces for the
C-ABI's subroutine entry points (include/g3m3.h).

Names, argument order and C types come from the header.  Whether a dummy is
an array or a scalar comes from the reference subroutine of the same name
(its dummy declarations under code/), so that a caller which ``use``s the
module passes exactly what it passed to the reference.  Entry points with no
reference counterpart take their array arguments from ``EXTRA_ARRAYS``.

Run once, in a tree where the reference sources are present:
    python -m g3m3_amd.fortran.make_iface /root/reference/code
The generated module is committed; building it needs only a Fortran
compiler (g3m3_amd/build.py compiles it with flang and checks a caller).
"""
from __future__ import annotations

import re
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
HEADER = ROOT / "include" / "g3m3.h"
OUT = Path(__file__).resolve().parent / "g3m3_iface.f90"

# entry points with no reference subroutine: their array (pointer) arguments
EXTRA_ARRAYS = {
    "bande_eta": {"efldx", "efldy", "efldz", "curx", "cury", "curz", "qrho", "hrho", "orho"},
    "bfield_total": {"bx", "by", "bz", "bx0", "by0", "bz0", "bsx", "bsy", "bsz", "btot"},
    "species_velocity": {"rho", "px", "py", "pz", "vx", "vy", "vz"},
    "divb_correct_ex": {"xspac", "bx", "by", "bz"},
    "sample_fields": {"sxyz", "grid_minvals", "grid_maxvals", "xspac", "fields", "scdata",
                      "gridpt"},
    "sample_fields_adjoint": {"sxyz", "grid_minvals", "grid_maxvals", "xspac", "fields_bar",
                              "scdata_bar"},
    "sample_craft": {"xcraft", "zcraft", "grid_minvals", "grid_maxvals", "xspac", "fields",
                     "sxyz", "scdata", "gridpt"},
    "push_ion_adjoint": {"bsx", "bsy", "bsz", "btot", "efldx", "efldy", "efldz", "tvx", "tvy",
                         "tvz", "bsx_bar", "bsy_bar", "bsz_bar", "btot_bar", "efldx_bar",
                         "efldy_bar", "efldz_bar", "tvx_bar", "tvy_bar", "tvz_bar", "work"},
    "lap_plasma_adjoint": {"vx", "vy", "vz", "vx_bar", "vy_bar", "vz_bar", "work"},
    "species_velocity_adjoint": {"rho", "px", "py", "pz", "vx_bar", "vy_bar", "vz_bar",
                                 "rho_bar", "px_bar", "py_bar", "pz_bar"},
}
# entry points with the same dummies as a reference subroutine
SAME_AS = {"set_speed_agrd_noalloc": "set_speed_agrd", "bndry_flanks_v": None,
           "bndry_corer_v": None}
# not bound here: handles, descriptors, host-side C strings and the run API
SKIP_PREFIX = ("run_", "batch_", "field_model_", "full_model_", "prof_", "fluid_", "init_plasma",
               "graphing_cuts", "torus_inject", "bndry_flanks_v", "bndry_corer_v", "set_stream", "last_error_msg")

# dummies the reference leaves implicitly typed (INTEGER by their initial
# letter) while its main program passes a LOGICAL (code/3d_main.f90 declares
# isotropic logical): bound as the LOGICAL the caller passes
LOGICAL_ALWAYS = {"isotropic"}

CTYPE = {"float": "real(c_float)", "int": "integer(c_int)", "double": "real(c_double)",
         "long long": "integer(c_long_long)"}


def header_decls():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = []
    for m in re.finditer(r"\b(void|int)\s+g3m3_([a-z0-9_]+)\s*\(([^;{]*?)\)\s*;", text, re.S):
        ret, name, args = m.group(1), m.group(2), m.group(3)
        if name.endswith("_") or name in [d[1] for d in out]:
            continue
        params = []
        a = " ".join(args.split())
        if a not in ("", "void"):
            for p in a.split(","):
                p = p.strip()
                mm = re.match(r"(const\s+)?(float|int|double|long long|char|void|size_t)"
                              r"\s*(\*?)\s*(const\s*)?(\*?)\s*([A-Za-z0-9_]+)$", p)
                if not mm:
                    raise ValueError(f"{name}: cannot read parameter '{p}'")
                const, ct, s1, _, s2, pn = mm.groups()
                params.append((pn, ct, (s1 + s2).count("*"), bool(const)))
        out.append((ret, name, params))
    return out


def ref_arrays(code: Path):
    """{subroutine name (lower case): set of its array dummies}."""
    res = {}
    for f in sorted(code.glob("*.f90")):
        src = f.read_text(errors="replace")
        lines = []
        for ln in src.splitlines():
            ln = ln.split("!")[0].rstrip()
            if not ln.strip() and lines and lines[-1].endswith("&"):
                continue            # blank or comment line inside a continuation
            if lines and lines[-1].endswith("&"):
                lines[-1] = lines[-1][:-1] + " " + ln.strip().lstrip("&")
            else:
                lines.append(ln.strip())
        cur = None
        for ln in lines:
            low = ln.lower()
            m = re.match(r"(?:recursive\s+)?subroutine\s+([a-z0-9_]+)\s*\((.*)\)", low)
            if m:
                cur = m.group(1)
                res[cur] = (set(a.strip() for a in m.group(2).split(",")), set(), set())
                continue
            if cur is None:
                continue
            if re.match(r"end\s*subroutine", low) or low == "end":
                cur = None
                continue
            if not re.match(r"(real|integer|logical|character|double|dimension|complex)\b", low):
                continue
            dummies, arrs, logs = res[cur]
            attr_dim = re.match(r"[^:]*\bdimension\s*\(", low) and "::" in low
            body = low.split("::", 1)[1] if "::" in low else re.sub(
                r"^(real|integer|logical|character|double\s+precision|dimension|complex)"
                r"(\s*\*\s*\d+|\s*\([^)]*\))?", "", low)
            depth, name, names = 0, "", []
            for ch in body + ",":
                if ch == "(":
                    depth += 1
                    if depth == 1:
                        names.append((name.strip(), True))
                        name = None
                    continue
                if ch == ")":
                    depth -= 1
                    continue
                if depth:
                    continue
                if ch == ",":
                    if name is not None and name.strip():
                        names.append((name.strip(), False))
                    name = ""
                    continue
                if name is not None:
                    name += ch
            for nm, isarr in names:
                nm = nm.split("=")[0].strip()
                if nm in dummies and (isarr or attr_dim):
                    arrs.add(nm)
                if nm in dummies and low.startswith("logical"):
                    logs.add(nm)
    return {k: v[1] for k, v in res.items()}, {k: v[2] for k, v in res.items()}


def emit(decls, arrays, logicals):
    o = ["! Generated by g3m3_amd/fortran/make_iface.py from include/g3m3.h; do not edit.",
         "! bind(c) interfaces of the g3m3 C-ABI's subroutine entry points.  Each",
         "! dummy is an array or a scalar as in the reference subroutine of the same",
         "! name, so `use g3m3_iface` then `call g3m3_<name>(<the reference's",
         "! arguments>)` is checked by the compiler.  The reference's LOGICAL dummies",
         "! are logical(c_int) (4 bytes, as the library reads them: nonzero = .true.).",
         "! Arrays are host arrays registered with g3m3_register or device arrays.",
         "module g3m3_iface",
         "    use, intrinsic :: iso_c_binding",
         "    implicit none",
         "    interface",
         "        integer(c_int) function g3m3_init() bind(c, name='g3m3_init')",
         "            import :: c_int",
         "        end function",
         "        integer(c_int) function g3m3_last_error() bind(c, name='g3m3_last_error')",
         "            import :: c_int",
         "        end function",
         "        subroutine g3m3_sync() bind(c, name='g3m3_sync')",
         "        end subroutine",
         "        integer(c_int) function g3m3_register(host, bytes) bind(c, name='g3m3_register')",
         "            import :: c_int, c_long_long",
         "            type(*), dimension(*) :: host",
         "            integer(c_long_long), value :: bytes",
         "        end function"]
    for fn in ("unregister", "to_device", "to_host"):
        o += [f"        integer(c_int) function g3m3_{fn}(host) bind(c, name='g3m3_{fn}')",
              "            import :: c_int",
              "            type(*), dimension(*) :: host",
              "        end function"]
    count = 0
    for ret, name, params in decls:
        if name in ("init", "last_error", "sync", "register", "unregister", "to_device",
                    "to_host") or name.startswith(SKIP_PREFIX):
            continue
        ref = SAME_AS.get(name, name)
        arr = arrays.get(ref) if ref else None
        if arr is None:
            arr = EXTRA_ARRAYS.get(name)
        if arr is None:
            raise ValueError(f"no array/scalar information for g3m3_{name}")
        names = [p[0] for p in params]
        kind = "integer(c_int) function" if ret == "int" else "subroutine"
        head = f"        {kind} g3m3_{name}("
        line, body = head, []
        for i, n in enumerate(names):
            piece = n + (", " if i + 1 < len(names) else "")
            if len(line) + len(piece) > 88:
                body.append(line.rstrip() + " &")
                line = "            "
            line += piece
        body.append(line + f") bind(c, name='g3m3_{name}')")
        kinds = set()
        decl = []
        for pn, ct, ptr, const in params:
            if ct == "size_t":
                decl.append(f"            integer(c_size_t), value :: {pn}")
                kinds.add("c_size_t")
                continue
            if ct == "char":
                decl.append(f"            character(kind=c_char), intent(in) :: {pn}(*)")
                kinds.add("c_char")
                continue
            if ptr == 2:
                decl.append(f"            type(c_ptr), intent(in) :: {pn}(*)")
                kinds.add("c_ptr")
                continue
            ft = CTYPE[ct]
            if ct == "int" and (pn.lower() in LOGICAL_ALWAYS or
                                (ref and pn.lower() in logicals.get(ref, set()))):
                ft = "logical(c_int)"
            kinds.add(ft.split("(")[1].rstrip(")"))
            intent = "in" if const else "inout"
            dim = "(*)" if pn.lower() in arr else ""
            decl.append(f"            {ft}, intent({intent}) :: {pn}{dim}")
        if ret == "int":
            kinds.add("c_int")
        o += body
        o.append("            import :: " + ", ".join(sorted(kinds)))
        o += decl
        o.append("        end " + ("function" if ret == "int" else "subroutine"))
        count += 1
    o += ["    end interface", "end module g3m3_iface", ""]
    return "\n".join(o), count


def main(argv):
    code = Path(argv[1]) if len(argv) > 1 else Path("/root/reference/code")
    arrays, logicals = ref_arrays(code)
    text, n = emit(header_decls(), arrays, logicals)
    OUT.write_text(text)
    print(f"wrote {OUT} ({n} entry points)")


if __name__ == "__main__":
    main(sys.argv)
