"""Argument lists of the C-ABI entry points, by name (include/g3m3.h).

Each builder returns the names of a subroutine's arguments in the
reference's dummy order; the driver resolves each name to a device tensor,
a small host array or a scalar.  ``pre`` is the prefix of the array set
being passed: "" the state, "old" the step-start copy, "wrk" the
half-step / smoothed work set.
"""
SPECIES = "qho"
GRID = ["grd_xmin", "grd_xmax", "grd_ymin", "grd_ymax", "grd_zmin", "grd_zmax"]
ION = ("rho", "presx", "presy", "presz", "px", "py", "pz", "presxy", "presxz", "presyz")
# the reference's bndry_flanks / bndry_corer field order
FLANK_ORDER = [s + f for s in SPECIES for f in ("rho", "px", "py", "pz", "presx", "presy",
                                                 "presz", "presxy", "presxz", "presyz")] + \
              ["epres", "bx", "by", "bz"]
CORER_ORDER = ([s + f for s in SPECIES for f in ("rho", "presx", "presy", "presz", "px", "py",
                                                  "pz")] + ["epres", "bx", "by", "bz"]
               + [s + f for s in SPECIES for f in ("presxy", "presxz", "presyz")])


def push_ion(out, old, src, sp, rmass):
    """code/sub_push_ion.f90:4-17"""
    def g(pre, fs):
        return [pre + sp + f for f in fs]
    d = ("rho", "presx", "presy", "presz", "px", "py", "pz")
    o = ("presxy", "presxz", "presyz")
    return (g(out, d) + g(old, d) + g(src, d) + g(out, o) + g(old, o) + g(src, o)
            + ["bsx", "bsy", "bsz", "btot", "efldx", "efldy", "efldz", rmass,
               "vvx", "vvy", "vvz", "tvx", "tvy", "tvz", "gamma", "gamma1",
               "nx", "ny", "nz", "n_grids", "box", "delt", "grav", "re_equiv", "reynolds"]
            + GRID + ["isotropic"])


def set_rho(pre):
    """code/sub_set_rho.f90:5-11"""
    n = []
    for s, m in zip(SPECIES, ("rmassq", "rmassh", "rmasso")):
        n += [pre + s + f for f in ("rho", "presx", "presy", "presz", "presxy", "presxz",
                                    "presyz")] + [m]
    return n + [pre + "epres", "nx", "ny", "nz", "n_grids", "box", "o_conc"]


def set_speed(pre):
    """code/sub_set_speed_agrd.f90:6-22"""
    n = ["nx", "ny", "nz", "n_grids", "box", "rmassq", "rmassh", "rmasso", "isotropic",
         "smallbit", "aniso_limit", "vlim", "qvlim", "hvlim", "ovlim", "qclim", "hclim",
         "oclim", "cslim", pre + "bx", pre + "by", pre + "bz", "bx0", "by0", "bz0"]
    for s in SPECIES:
        n += [pre + s + f for f in ("rho", "px", "py", "pz", "presx", "presy", "presz",
                                    "presxy", "presxz", "presyz")]
    return n + [pre + "epres", "btot", "pxmax", "pymax", "pzmax", "csmax", "alfmax", "fastest"]


def bndry_outer(pre):
    """code/sub_bndry_outer.f90:5-16"""
    n = []
    for s in SPECIES:
        n += [pre + s + f for f in ("rho", "presx", "presy", "presz", "px", "py", "pz")]
    n += [pre + "epres"]
    for s in SPECIES:
        n += [pre + s + f for f in ("presxy", "presxz", "presyz")]
    return n + ["rmassq", "rmassh", "rmasso", pre + "bx", pre + "by", pre + "bz",
                "bx0", "by0", "bz0", "nx", "ny", "nz", "n_grids", "srho", "rho_frac",
                "o_conc", "spress", "spx", "spy", "spz", "sbx_wind", "sby_wind", "sbz_wind",
                "ti_te", "isotropic"]


def bndry_inner(pre):
    """code/sub_bndry_inner.f90:5-19, plus v_rot and r_rot (common /rotation/)"""
    n = []
    for s, m in zip(SPECIES, ("rmassq", "rmassh", "rmasso")):
        n += [pre + s + f for f in ("rho", "presx", "presy", "presz", "px", "py", "pz")] + [m]
    n += [pre + "epres", pre + "bx", pre + "by", pre + "bz"]
    for s in SPECIES:
        n += [pre + s + f for f in ("presxy", "presxz", "presyz")]
    return n + ["nx", "ny", "nz", "n_grids", "parm_srf", "parm_mid", "parm_zero",
                "ijsrf", "numsrf", "ijmid", "nummid", "ijzero", "numzero",
                "mbndry", "msrf", "mmid", "mzero", "erho", "epress", "alpha_e", "ti_te",
                "o_conc", "re_equiv", "r_inner", "sbx_wind", "spress"] + GRID + ["v_rot", "r_rot"]


def lap_plasma(sp, src, dst, v=("vvx", "vvy", "vvz")):
    """code/sub_lap_plasma.f90:4-10"""
    f = ("rho", "px", "py", "pz", "presx", "presy", "presz", "presxy", "presxz", "presyz")
    return ([src + sp + x for x in f] + [dst + sp + x for x in f] + list(v)
            + ["nx", "ny", "nz", "n_grids", "box", "chirho", "chipxyz", "chierg", "delt",
               "isotropic"])


def flux_correct():
    """code/sub_flux_correct.f90:4-35"""
    f = ("rho", "presx", "presy", "presz", "presxy", "presxz", "presyz", "px", "py", "pz")
    n = []
    for s in SPECIES:
        n += [s + x for x in f] + ["wrk" + s + x for x in f] + ["old" + s + x for x in f]
    return n + ["epres", "wrkepres", "oldepres", "bx", "by", "bz", "wrkbx", "wrkby", "wrkbz",
                "oldbx", "oldby", "oldbz", "vvx", "vvy", "vvz", "nx", "ny", "nz", "n_grids",
                "box", "difrho", "diferg", "xspac", "isotropic"]


def bande(pre, eta=False):
    """code/sub_BandE.f90:5-10 (+ eta for the uniform-resistivity variant)"""
    n = ["efldx", "efldy", "efldz", "bsx", "bsy", "bsz", "curx", "cury", "curz",
         "evx", "evy", "evz", "btot", pre + "epres", pre + "qrho", pre + "hrho", pre + "orho",
         "rst", "resist", "reynolds", "nx", "ny", "nz", "n_grids", "box", "rmassq", "rmassh",
         "rmasso", "ijmid", "nummid", "ijzero", "mbndry", "numzero", "mmid", "mzero",
         "rx", "ry", "rz"]
    return n + (["eta"] if eta else [])


def fnd_evel(pre):
    """code/sub_fnd_evel.f90:4-7"""
    return ([pre + "qpx", pre + "qpy", pre + "qpz", pre + "qrho", pre + "hpx", pre + "hpy",
             pre + "hpz", pre + "hrho", pre + "opx", pre + "opy", pre + "opz", pre + "orho",
             "curx", "cury", "curz", "evx", "evy", "evz", "tvx", "tvy", "tvz",
             "nx", "ny", "nz", "n_grids", "box", "rmassq", "rmassh", "rmasso", "reynolds"])
