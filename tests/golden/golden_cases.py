"""Seeded input cases for the golden fixtures: (subroutine, argument names,
arrays it writes, extra setup).  Shared by make_golden.py and the test."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from common import (World, push_ion_names, set_rho_names, set_speed_names,  # noqa: E402
                    bndry_outer_names, bndry_inner_names, lap_plasma_names,
                    flux_correct_names, flanks_names, corer_names, GRID)

F = np.float32


def _nothing(w, lib=None):
    pass


def _stress(w, lib=None):
    r = np.random.default_rng(77)
    for f in ("qrho", "hrho", "orho"):
        m = r.random(w.a[f].shape) < 0.1
        w.a[f][m] = r.uniform(-1e-4, 1e-4, m.sum()).astype(F)
    for f in ("qpresx", "hpresy", "opresz", "epres", "qpx", "hpy", "opz"):
        m = r.random(w.a[f].shape) < 0.1
        w.a[f][m] *= F(-25)


def _wind(w, lib=None):
    w.s.update(srho=F(25.), spress=F(0.02395), spx=F(10.), spy=F(0.5), spz=F(-0.1),
               sbx_wind=F(0.), sby_wind=F(0.01), sbz_wind=F(-0.244))


def _inner(w, lib=None):
    w.s.update(erho=F(1.), epress=F(1.), alpha_e=F(6.), r_inner=F(3.), sbx_wind=F(0.),
               spress=F(0.02))
    if lib is not None:
        rot = lib.rotation()
        rot.v_rot, rot.r_rot = 0.3, 2.0


def _times(w, lib=None):
    from oracle.ref import F64
    w.a["t_old"] = np.array([0.0, 0.1, 0.1], dtype=np.float64)
    w.a["t_new"] = np.array([0.05, 0.3, 0.5], dtype=np.float64)
    w.s["t"] = F64(0.17) if lib is not None else 0.17


SP = ["q" + f for f in ("rho", "presx", "presy", "presz", "px", "py", "pz", "presxy",
                         "presxz", "presyz")]
CASES = {
    "calcur": ("calcur", ["bx", "by", "bz", "nx", "ny", "nz", "n_grids", "box", "curx", "cury",
                          "curz", "rx", "ry", "rz"], ["curx", "cury", "curz"], _nothing),
    "bande": ("BandE", ["efldx", "efldy", "efldz", "bsx", "bsy", "bsz", "curx", "cury", "curz",
                        "evx", "evy", "evz", "btot", "epres", "qrho", "hrho", "orho", "rst",
                        "resist", "reynolds", "nx", "ny", "nz", "n_grids", "box", "rmassq",
                        "rmassh", "rmasso", "ijmid", "nummid", "ijzero", "mbndry", "numzero",
                        "mmid", "mzero", "rx", "ry", "rz"], ["efldx", "efldy", "efldz"],
              _nothing),
    "push_ion": ("push_ion", push_ion_names("", "old", "wrk", "q", "rmassq"), SP, _nothing),
    "set_rho": ("set_rho", set_rho_names(""), SP + ["epres"], _stress),
    "set_speed_agrd": ("set_speed_agrd", set_speed_names(""),
                       SP + ["epres", "btot", "fastest"], _stress),
    "lap_plasma": ("lap_plasma", lap_plasma_names("q", "", "wrk"), ["wrk" + k for k in SP],
                   _nothing),
    "flux_correct": ("flux_correct", flux_correct_names(), SP + ["epres", "bx", "by", "bz"],
                     _nothing),
    "bndry_outer": ("bndry_outer", bndry_outer_names(""), SP + ["epres", "bx", "by", "bz"],
                    _wind),
    "bndry_inner": ("bndry_inner", bndry_inner_names(""), SP + ["epres", "bx", "by", "bz"],
                    _inner),
    "bndry_flanks": ("bndry_flanks", flanks_names("wrk", "", "old"), ["wrk" + k for k in SP],
                     _times),
    "bndry_corer": ("bndry_corer", corer_names(""), SP, _nothing),
}


def world_for(cid):
    if cid in ("bndry_flanks", "bndry_corer"):
        w = World(13, 13, 7, 3, seed=7, lists=False)
        w.s["box"] = 1
    else:
        w = World(11, 9, 8, 2, seed=11)
        w.s.update(box=2 if cid != "bndry_inner" else 1, delt=F(0.037), vlim=F(1.2),
                   rx=F(2.), ry=F(2.), rz=F(2.))
    return w
