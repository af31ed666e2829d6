"""write_graphing_data (code/sub_write_graphing_data.f90) around the device's
cut-plane tables.

For every box the device computes, on three cut planes (xy, xz, yz), the
quantities the reference tabulates for plotting (``g3m3_graphing_cuts``:
the whole-box calcur, total B, fnd_evel, BandE, species velocities and
fnd_pres, then the physical-unit columns).  This module writes them as the
reference does: figures/data/<run_name>_{plas,flow,pres,bande,model}_
<box><cut>_t<nnn>.dat, the same headers, column titles and Fortran edit
descriptors (ES14.6; F9.2 for the model file's positions).  The plot
number nnn is one more than the number of <run_name>_t???_data.tar.gz
archives already in figures/data (:235-250).  The reference then starts
its Python plotting script (:818-828); that call stays with the caller.
"""
from __future__ import annotations

import glob
import math
import os

import numpy as np

from . import _lib

F = np.float32
CUTS = ("xy", "xz", "yz")
NCOL = {"plas": 8, "flow": 13, "pres": 10, "bande": 11, "model": 11}
ORDER = ("plas", "flow", "pres", "bande", "model")
TITLES = {
    "plas": ("log(qden/cc)", "qtemp(eV)", "log(hden/cc)", "htemp(eV)", "log(oden/cc)",
             "otemp(eV)", "log(eden/cc)", "etemp(eV)"),
    "flow": ("qvx(km/s)", "qvy(km/s)", "qvz(km/s)", "hvx(km/s)", "hvy(km/s)", "hvz(km/s)",
             "ovx(km/s)", "ovy(km/s)", "ovz(km/s)", "u_netx(km/s)", "u_nety(km/s)",
             "u_netz(km/s)", "|v|/V_Alf"),
    "pres": ("log_qpar(nPa)", "log_qprp(nPa)", "log_qcrs(nPa)", "log_hpar(nPa)",
             "log_hprp(nPa)", "log_hcrs(nPa)", "log_opar(nPa)", "log_oprp(nPa)",
             "log_ocrs(nPa)", "log_eprs(nPa)"),
    "bande": ("Bx(nT)", "By(nT)", "Bz(nT)", "|Btot|(nT)", "Ex(mV/m)", "Ey(mV/m)", "Ez(mV/m)",
              "|Etot|(mV/m)", "Jx(nA/m^2)", "Jy(nA/m^2)", "Jz(nA/m^2)"),
    "model": ("x(R_E)", "y(R_E)", "z(R_E)", "rad(R_E)", "Bdx(nT)", "Bdy(nT)", "Bdz(nT)",
              "Bdipx(nT)", "Bdipy(nT)", "Bdipz(nT)", "|Bdip|(nT)"),
}
DESCRIBE = {
    "plas": "  Plasma densities and temperatures for each species, in log(1/cm^3) and eV.",
    "flow": "  Flow velocity components for each species and net flow, in km/s, and Alfven "
            "Mach number.",
    "pres": "  Pressure tensor terms for each species: parallel, perpendicular, and cross, and "
            "isotropic electron pressure, all in log(nPa).",
    "bande": "  Net magnetic and electric field components and magnitudes, and current density "
             "components, in nT, mV/m, and nA/m^2,.",
    "model": "  Planetocentric distance in planetary radii, magnetic field comps for perturbed "
             "and background, and BG magnitude, all in nT.",
}
INTRO = "This file contains data for the following quantities:"


def _es(x: float, w: int = 14, d: int = 6) -> str:
    """Fortran ESw.d of a single-precision value."""
    if math.isnan(x):
        return "NaN".rjust(w)
    if math.isinf(x):
        return ("Inf" if x > 0 else "-Inf").rjust(w)
    return f"{x:{w}.{d}E}"


def _f(x: float, w: int = 9, d: int = 2) -> str:
    if math.isnan(x):
        return "NaN".rjust(w)
    return f"{x:{w}.{d}f}"


def cut_tables(run, box: int, limit: float, re_equiv, b_equiv, v_equiv, rho_equiv,
               planet_rad):
    """Device tables of one box: a list of three (npts, 53) float32 arrays,
    one per cut plane, in the reference's point order."""
    import torch
    r = run
    if hasattr(r, "sync_arrays"):
        r.sync_arrays()
    nx, ny, nz = r.nx, r.ny, r.nz
    n = nx * ny + nx * nz + ny * nz
    out = torch.zeros(n * 53, dtype=torch.float32, device="cuda")
    g = r.wl.grid
    gmin = np.ascontiguousarray(np.stack([g["grd_xmin"], g["grd_ymin"], g["grd_zmin"]], 1), F)
    s = r.s
    sp = []
    for q in "qho":
        sp += [r.a[q + k] for k in ("rho", "px", "py", "pz", "presx", "presy", "presz")]
    _lib.call("graphing_cuts", *sp, r.a["bx"], r.a["by"], r.a["bz"], r.a["epres"], r.a["bx0"],
              r.a["by0"], r.a["bz0"], r.a["rst"], r.a["ijzero"], r.a["numzero"], r.a["ijmid"],
              r.a["nummid"], nx, ny, nz, r.ngrd, box, int(s["mbndry"]), int(s["mmid"]),
              int(s["mzero"]), F(limit), r.a["xspac"], gmin, F(s["rmassq"]), F(s["rmassh"]),
              F(s["rmasso"]), F(s["reynolds"]), F(s["resist"]), F(s["smallbit"]), F(re_equiv),
              F(b_equiv), F(v_equiv), F(rho_equiv), F(planet_rad), out)
    _lib.sync()
    h = out.cpu().numpy().reshape(n, 53)
    return [h[:nx * ny], h[nx * ny:nx * ny + nx * nz], h[nx * ny + nx * nz:]]


def write_graphing_data(run, out_dir: str, run_name: str, ut: float, limit: float,
                        re_equiv, b_equiv, v_equiv, rho_equiv, planet_rad) -> int:
    """The reference's graphing files for every box of the run under
    out_dir/figures/data; returns the plot number (nplots)."""
    data = os.path.join(out_dir, "figures", "data")
    os.makedirs(data, exist_ok=True)
    name8 = run_name[:8].ljust(8)          # character*8
    start = os.path.join(data, run_name.strip() + "_")
    nplots = len(glob.glob(start + "t???_data.tar.gz")) + 1
    ut_string = f"{ut:9.4f} "              # F9.4 into character*10
    for box in range(1, run.ngrd + 1):
        tabs = cut_tables(run, box, limit, re_equiv, b_equiv, v_equiv, rho_equiv, planet_rad)
        for m, cut in enumerate(CUTS):
            t = tabs[m].astype(np.float64)
            col = 0
            for kind in ORDER:
                nc = NCOL[kind]
                block = t[:, col:col + nc]
                col += nc
                path = f"{start}{kind}_{box}{cut}_t{nplots:03d}.dat"
                head = "\n".join([INTRO, DESCRIBE[kind], "  run name:      " + name8,
                                  f"  box, cut:      {box}, {cut}-plane",
                                  f"  ut, #in seq: {ut_string}, {nplots:03d}"])
                if kind == "model":
                    titles = "".join(x.rjust(9) for x in TITLES[kind][:4]) + \
                        "".join(x.rjust(14) for x in TITLES[kind][4:])
                else:
                    titles = "".join(x.rjust(14) for x in TITLES[kind])
                with open(path, "w") as f:
                    f.write(head.rstrip() + "\n")
                    f.write(titles + "\n")
                    if kind == "model":
                        for row in block:
                            f.write("".join(_f(v) for v in row[:4])
                                    + "".join(_es(v) for v in row[4:]) + "\n")
                    else:
                        for row in block:
                            f.write("".join(_es(v) for v in row) + "\n")
    return nplots
