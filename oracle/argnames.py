"""Argument lists of the reference's hot-path subroutines, by name.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Each builder returns
the names of a subroutine's dummies in the reference's order; ``pre`` is
the array-name prefix of the set being passed ("" state, "old", "wrk").
A name resolves in a run's dict of arrays or scalars, so one list serves
the reference build, the NumPy restatement and the HIP C-ABI alike.
"""

SPECIES = "qho"
GRID = ["grd_xmin", "grd_xmax", "grd_ymin", "grd_ymax", "grd_zmin", "grd_zmax"]


def push_ion_names(out, old, wrk, sp, rmass):
    """Argument names of code/sub_push_ion.f90:4-17."""
    def g(pre, names):
        return [pre + sp + n for n in names]
    return (g(out, ["rho", "presx", "presy", "presz", "px", "py", "pz"])
            + g(old, ["rho", "presx", "presy", "presz", "px", "py", "pz"])
            + g(wrk, ["rho", "presx", "presy", "presz", "px", "py", "pz"])
            + g(out, ["presxy", "presxz", "presyz"])
            + g(old, ["presxy", "presxz", "presyz"])
            + g(wrk, ["presxy", "presxz", "presyz"])
            + ["bsx", "bsy", "bsz", "btot", "efldx", "efldy", "efldz", rmass,
               "vvx", "vvy", "vvz", "tvx", "tvy", "tvz", "gamma", "gamma1",
               "nx", "ny", "nz", "n_grids", "box", "delt", "grav", "re_equiv", "reynolds"]
            + GRID + ["isotropic"])


def set_rho_names(pre):
    """code/sub_set_rho.f90:5-11."""
    n = []
    for s, m in zip(SPECIES, ("rmassq", "rmassh", "rmasso")):
        n += [pre + s + f for f in ("rho", "presx", "presy", "presz", "presxy", "presxz",
                                    "presyz")] + [m]
    return n + [pre + "epres", "nx", "ny", "nz", "n_grids", "box", "o_conc"]


def set_speed_names(pre):
    """code/sub_set_speed_agrd.f90:6-22."""
    n = ["nx", "ny", "nz", "n_grids", "box", "rmassq", "rmassh", "rmasso", "isotropic",
         "smallbit", "aniso_limit", "vlim", "qvlim", "hvlim", "ovlim", "qclim", "hclim",
         "oclim", "cslim", pre + "bx", pre + "by", pre + "bz", "bx0", "by0", "bz0"]
    for s in SPECIES:
        n += [pre + s + f for f in ("rho", "px", "py", "pz", "presx", "presy", "presz",
                                    "presxy", "presxz", "presyz")]
    return n + [pre + "epres", "btot", "pxmax", "pymax", "pzmax", "csmax", "alfmax",
                "fastest"]


def bndry_outer_names(pre):
    """code/sub_bndry_outer.f90:5-16."""
    n = []
    for s in SPECIES:
        n += [pre + s + f for f in ("rho", "presx", "presy", "presz", "px", "py", "pz")]
    n += [pre + "epres"]
    for s in SPECIES:
        n += [pre + s + f for f in ("presxy", "presxz", "presyz")]
    n += ["rmassq", "rmassh", "rmasso", pre + "bx", pre + "by", pre + "bz",
          "bx0", "by0", "bz0", "nx", "ny", "nz", "n_grids",
          "srho", "rho_frac", "o_conc", "spress", "spx", "spy", "spz",
          "sbx_wind", "sby_wind", "sbz_wind", "ti_te", "isotropic"]
    return n


def bndry_inner_names(pre):
    """code/sub_bndry_inner.f90:5-19."""
    n = []
    for s, m in zip(SPECIES, ("rmassq", "rmassh", "rmasso")):
        n += [pre + s + f for f in ("rho", "presx", "presy", "presz", "px", "py", "pz")] + [m]
    n += [pre + "epres", pre + "bx", pre + "by", pre + "bz"]
    for s in SPECIES:
        n += [pre + s + f for f in ("presxy", "presxz", "presyz")]
    n += ["nx", "ny", "nz", "n_grids", "parm_srf", "parm_mid", "parm_zero",
          "ijsrf", "numsrf", "ijmid", "nummid", "ijzero", "numzero",
          "mbndry", "msrf", "mmid", "mzero", "erho", "epress", "alpha_e", "ti_te",
          "o_conc", "re_equiv", "r_inner", "sbx_wind", "spress"] + GRID
    return n


def lap_plasma_names(sp, src, dst):
    """code/sub_lap_plasma.f90:4-10 (v in vvx..vvz)."""
    f = ("rho", "px", "py", "pz", "presx", "presy", "presz", "presxy", "presxz", "presyz")
    return ([src + sp + x for x in f] + [dst + sp + x for x in f]
            + ["vvx", "vvy", "vvz", "nx", "ny", "nz", "n_grids", "box",
               "chirho", "chipxyz", "chierg", "delt", "isotropic"])


def flux_correct_names():
    """code/sub_flux_correct.f90:4-35."""
    f = ("rho", "presx", "presy", "presz", "presxy", "presxz", "presyz", "px", "py", "pz")
    n = []
    for s in SPECIES:
        n += [s + x for x in f] + ["wrk" + s + x for x in f] + ["old" + s + x for x in f]
    n += ["epres", "wrkepres", "oldepres", "bx", "by", "bz", "wrkbx", "wrkby", "wrkbz",
          "oldbx", "oldby", "oldbz", "vvx", "vvy", "vvz", "nx", "ny", "nz", "n_grids", "box",
          "difrho", "diferg", "xspac", "isotropic"]
    return n


def flanks_names(tgt, cur, old):
    """code/sub_bndry_flanks.f90:5-35 argument layout."""
    order = []
    for s in SPECIES:
        order += [s + x for x in ("rho", "px", "py", "pz", "presx", "presy", "presz",
                                  "presxy", "presxz", "presyz")]
    order += ["epres", "bx", "by", "bz"]
    return ([tgt + x for x in order] + [cur + x for x in order] + [old + x for x in order]
            + ["work", "nx", "ny", "nz", "n_grids", "box", "t_old", "t_new", "t"] + GRID)


def corer_names(pre):
    """code/sub_bndry_corer.f90:5-14."""
    n = []
    for s in SPECIES:
        n += [pre + s + f for f in ("rho", "presx", "presy", "presz", "px", "py", "pz")]
    n += [pre + "epres", pre + "bx", pre + "by", pre + "bz"]
    for s in SPECIES:
        n += [pre + s + f for f in ("presxy", "presxz", "presyz")]
    return n + ["nx", "ny", "nz", "n_grids", "box"] + GRID
