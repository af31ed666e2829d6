"""Shared fixtures for the parity tests: seeded synthetic inputs laid out
exactly as the reference's Fortran arrays, and the reference argument
lists of each hot-path subroutine, so one list can be handed to the
reference build, to the oracle and to the HIP C-ABI."""
from __future__ import annotations

import copy

import numpy as np

F = np.float32
SPECIES = "qho"
SP_FIELDS = ("rho", "presx", "presy", "presz", "px", "py", "pz",
             "presxy", "presxz", "presyz")
# the 34 per-cell state fields
FIELDS = [s + f for s in SPECIES for f in SP_FIELDS] + ["epres", "bx", "by", "bz"]


from oracle.argnames import (GRID, push_ion_names, set_rho_names,  # noqa: F401,E402
                             set_speed_names, bndry_outer_names, bndry_inner_names,
                             lap_plasma_names, flux_correct_names, flanks_names,
                             corer_names)


class World:
    """Named arrays + scalars of one synthetic run."""

    def __init__(self, nx, ny, nz, ngrd, seed=0, nested=None, lists=True):
        self.nx, self.ny, self.nz, self.ngrd = nx, ny, nz, ngrd
        self.rng = np.random.default_rng(seed)
        self.a = {}
        self.s = dict(nx=nx, ny=ny, nz=nz, n_grids=ngrd, rmassq=F(1.), rmassh=F(1.),
                      rmasso=F(16.), gamma=F(1.6666), gamma1=F(0.6666), reynolds=F(256.),
                      resist=F(64.), grav=F(0.05), re_equiv=F(0.2), isotropic=False,
                      smallbit=F(1e-5), aniso_limit=F(0.9995), qvlim=F(1.), hvlim=F(1.4),
                      ovlim=F(2.), qclim=F(1.), hclim=F(1.), oclim=F(4.), cslim=F(1.),
                      chirho=F(2.), chipxyz=F(2.), chierg=F(0.), difrho=F(0.008),
                      diferg=F(0.004), o_conc=F(0.125), rho_frac=F(0.001), ti_te=F(5.),
                      mbndry=1)
        # grid extents: box b spacing 2^(b-1), centred on the origin
        L = (nx - 1) // 2
        xs = np.array([2.0 ** b for b in range(ngrd)], dtype=F)
        self.a["xspac"] = xs
        self.a["grd_xmin"] = (-(nx - 1) / 2 * xs).astype(F)
        self.a["grd_xmax"] = ((nx - 1) / 2 * xs).astype(F)
        self.a["grd_ymin"] = (-(ny - 1) / 2 * xs).astype(F)
        self.a["grd_ymax"] = ((ny - 1) / 2 * xs).astype(F)
        self.a["grd_zmin"] = (-(nz - 1) / 2 * xs).astype(F)
        self.a["grd_zmax"] = ((nz - 1) / 2 * xs).astype(F)
        del L
        shp4 = (ngrd, nz, ny, nx)
        r = self.rng
        for pre in ("", "old", "wrk"):
            for s in SPECIES:
                self.a[pre + s + "rho"] = r.uniform(0.5, 2.0, shp4).astype(F)
                for c in "xyz":
                    self.a[pre + s + "p" + c] = r.uniform(-0.3, 0.3, shp4).astype(F)
                    self.a[pre + s + "pres" + c] = r.uniform(0.1, 1.0, shp4).astype(F)
                for c in ("xy", "xz", "yz"):
                    self.a[pre + s + "pres" + c] = r.uniform(-0.05, 0.05, shp4).astype(F)
            self.a[pre + "epres"] = r.uniform(0.1, 0.5, shp4).astype(F)
            for c in "xyz":
                self.a[pre + "b" + c] = r.uniform(-0.3, 0.3, shp4).astype(F)
        for c in "xyz":
            self.a["b" + c + "0"] = r.uniform(-1.0, 1.0, shp4).astype(F)
        shp3 = (nz, ny, nx)
        for nm in ("curx", "cury", "curz", "evx", "evy", "evz", "tvx", "tvy", "tvz",
                   "efldx", "efldy", "efldz", "bsx", "bsy", "bsz", "btot",
                   "vvx", "vvy", "vvz", "work"):
            self.a[nm] = r.uniform(-1, 1, shp3).astype(F)
        for nm in ("bsx", "bsy", "bsz", "evx", "evy", "evz", "tvx", "tvy", "tvz",
                   "efldx", "efldy", "efldz"):
            self.a[nm] = r.uniform(-0.5, 0.5, shp3).astype(F)
        self.a["btot"] = np.maximum(np.sqrt(self.a["bsx"] ** 2 + self.a["bsy"] ** 2
                                            + self.a["bsz"] ** 2), F(1e-5)).astype(F)
        self.a["rst"] = r.uniform(0.0, 0.01, (1, nz, ny, nx)).astype(F)
        if lists:
            self._lists()
        # one-element outputs
        for nm in ("pxmax", "pymax", "pzmax", "csmax", "alfmax", "fastest",
                   "residual2", "scdata"):
            self.a[nm] = np.zeros(1, F)

    def _lists(self):
        """Inner-boundary index lists on box 1: spherical shells about the
        box centre (the layout code/3d_main.f90:1383-1455 builds)."""
        nx, ny, nz = self.nx, self.ny, self.nz
        k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        r = np.sqrt((i - (nx - 1) / 2) ** 2 + (j - (ny - 1) / 2) ** 2 + (k - (nz - 1) / 2) ** 2)
        rin = min(nx, ny, nz) / 4.0
        zero = r < rin - 1.5
        mid = (r >= rin - 1.5) & (r < rin - 0.5)
        srf = (r >= rin - 0.5) & (r <= rin + 0.6)
        for nm, m in (("srf", srf), ("mid", mid), ("zero", zero)):
            kk, jj, ii = np.nonzero(m)
            n = len(ii)
            lst = np.zeros((max(n, 1), 3, 1), dtype=np.int32)
            lst[:n, 0, 0] = ii + 1
            lst[:n, 1, 0] = jj + 1
            lst[:n, 2, 0] = kk + 1
            self.a["ij" + nm] = lst
            self.a["num" + nm] = np.array([n], dtype=np.int32)
            self.s["m" + nm] = max(n, 1)
            parm = np.zeros((max(n, 1), 7, 1), dtype=F)
            parm[:n, 0:3, 0] = self.rng.uniform(0.5, 2.0, (n, 3))
            parm[:n, 3:7, 0] = self.rng.uniform(0.1, 1.0, (n, 4))
            self.a["parm_" + nm] = parm

    def copy(self):
        w = copy.copy(self)
        w.a = {k: v.copy() for k, v in self.a.items()}
        w.s = dict(self.s)
        return w

    def args(self, names):
        out = []
        for n in names:
            if isinstance(n, str):
                if n in self.a:
                    out.append(self.a[n])
                elif n in self.s:
                    out.append(self.s[n])
                else:
                    raise KeyError(n)
            else:
                out.append(n)
        return out


def max_rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), 1e-30)
    return float(np.abs(a - b).max() / scale)


# --- the initialisation block (code/3d_main.f90:1238-1633) ----------------------

INIT_SCALARS = ("rmassq", "rmassh", "rmasso", "ti_te", "re_equiv", "r_inner", "zero_bndry",
                "mid_bndry", "srf_bndry", "alpha_e", "erho", "d_min", "cs_inner", "srho", "spx",
                "spy", "spz", "spress", "cs_wind", "gamma", "rho_frac", "resist", "v_rot",
                "r_rot", "rot_angle", "xdip", "ydip", "zdip", "sin_tilt", "cos_tilt", "b0")


def init_case(nx=21, ny=19, nz=13, ngrd=2, erho=0.0, mcap=400):
    """Inputs of the initialisation block for nested boxes centred on the
    planet, with corotation, a tilted and rotated dipole and a solar wind
    whose inner edge (r_rot/re_equiv = 6) lies inside the outer box."""
    s = dict(rmassq=1.0, rmassh=1.0, rmasso=16.0, ti_te=1.0, re_equiv=0.2, r_inner=3.0,
             zero_bndry=1.5, mid_bndry=2.5, srf_bndry=3.6, alpha_e=6.0, erho=erho, d_min=0.01,
             cs_inner=0.05, srho=0.5, spx=0.505, spy=0.0, spz=0.01, spress=0.02, cs_wind=0.3,
             gamma=1.6666, rho_frac=0.001, resist=64.0, v_rot=0.02, r_rot=1.2, rot_angle=0.3,
             xdip=0.00001, ydip=0.0, zdip=0.0, sin_tilt=float(F(np.sin(0.2))),
             cos_tilt=float(F(np.cos(0.2))), b0=270.0)
    s = {k: float(F(v)) for k, v in s.items()}
    xs = np.array([2.0 ** b for b in range(ngrd)], F)
    gmin = np.stack([-(nx - 1) / 2 * xs, -(ny - 1) / 2 * xs, -(nz - 1) / 2 * xs], 1).astype(F)
    gmax = np.stack([(nx - 1) / 2 * xs, (ny - 1) / 2 * xs, (nz - 1) / 2 * xs], 1).astype(F)
    shp = (ngrd, nz, ny, nx)
    f = {k: np.full(shp, 7.0, F) for k in FIELDS}
    b0 = [np.full(shp, 7.0, F) for _ in range(3)]
    mb = 1
    rst = np.full((mb, nz, ny, nx), 7.0, F)
    lists = dict(parm_srf=np.zeros((mcap, 7, mb), F), parm_mid=np.zeros((mcap, 7, mb), F),
                 parm_zero=np.zeros((mcap, 7, mb), F), ijsrf=np.zeros((mcap, 3, mb), np.int32),
                 ijmid=np.zeros((mcap, 3, mb), np.int32), ijzero=np.zeros((mcap, 3, mb), np.int32),
                 numsrf=np.zeros(mb, np.int32), nummid=np.zeros(mb, np.int32),
                 numzero=np.zeros(mb, np.int32))
    return dict(nx=nx, ny=ny, nz=nz, ngrd=ngrd, mbndry=mb, mcap=mcap, s=s, gmin=gmin, gmax=gmax,
                f=f, b0=b0, rst=rst, lists=lists)


def init_oracle(c):
    """Runs the oracle's init_plasma on a copy of an init_case; returns the
    copy and the list lengths found."""
    from oracle import restate as O
    d = copy.deepcopy(c)
    L = d["lists"]
    found = O.init_plasma(d["f"], *d["b0"], d["rst"], d["nx"], d["ny"], d["nz"], d["ngrd"],
                          d["mbndry"], d["mcap"], d["mcap"], d["mcap"], d["gmin"], d["gmax"],
                          L["parm_srf"], L["parm_mid"], L["parm_zero"], L["ijsrf"], L["ijmid"],
                          L["ijzero"], L["numsrf"], L["nummid"], L["numzero"], d["s"])
    return d, found


# --- the torus source (code/3d_main.f90:3149-3253) --------------------------------

TORUS_SCALARS = ("xdip", "ydip", "zdip", "re_equiv", "v_rot", "torus_inj_dist", "torus_rad",
                 "srf_bndry", "denq_torus", "denh_torus", "deno_torus", "rmassq", "rmassh",
                 "rmasso", "t_torus", "aniso_factor", "reduct")


def torus_case(nx=41, ny=41, nz=21, ngrd=2, seed=5, tilt=0.0):
    """A box of spacing 0.5 around the planet, B along z (tilted by `tilt`
    in x), and a torus of radius 5 and half-width 1 in the equatorial
    plane, with corotation."""
    rng = np.random.default_rng(seed)
    s = dict(xdip=0.0, ydip=0.0, zdip=0.0, re_equiv=1.0, v_rot=0.1, torus_inj_dist=5.0,
             torus_rad=1.0, srf_bndry=2.0, denq_torus=0.3, denh_torus=2.0, deno_torus=0.7,
             rmassq=1.0, rmassh=1.0, rmasso=16.0, t_torus=0.5, aniso_factor=0.8, reduct=0.9)
    s = {k: float(F(v)) for k, v in s.items()}
    xs = np.array([0.5 * 2.0 ** b for b in range(ngrd)], F)
    gmin = np.stack([-(nx - 1) / 2 * xs, -(ny - 1) / 2 * xs, -(nz - 1) / 2 * xs], 1).astype(F)
    gmax = np.stack([(nx - 1) / 2 * xs, (ny - 1) / 2 * xs, (nz - 1) / 2 * xs], 1).astype(F)
    shp = (ngrd, nz, ny, nx)
    f = {k: rng.uniform(0.5, 1.5, shp).astype(F) for k in FIELDS}
    for k in ("bx", "by", "bz"):
        f[k] = np.zeros(shp, F)
    b0 = [np.full(shp, F(tilt)), np.zeros(shp, F), np.ones(shp, F)]
    return dict(nx=nx, ny=ny, nz=nz, ngrd=ngrd, s=s, gmin=gmin, gmax=gmax, f=f, b0=b0)


def torus_oracle(c):
    from oracle import restate as O
    d = copy.deepcopy(c)
    tot = O.torus_inject(d["f"], *d["b0"], d["nx"], d["ny"], d["nz"], d["gmin"], d["gmax"],
                         d["s"])
    return d, tot


def same_bits(a, b) -> bool:
    """True when two arrays hold the same bits: float32 arrays compared as
    uint32 views (so -0.0 and 0.0 differ, and NaNs compare by payload),
    any other dtype element for element with the dtypes equal."""
    a, b = np.asarray(a), np.asarray(b)
    if a.shape != b.shape:
        return False
    if a.dtype == F and b.dtype == F:
        return bool(np.array_equal(np.ascontiguousarray(a).view(np.uint32),
                                   np.ascontiguousarray(b).view(np.uint32)))
    return a.dtype == b.dtype and bool(np.array_equal(a, b))


def bit_diffs(a, b) -> int:
    """Number of elements whose float32 bit patterns differ."""
    a = np.ascontiguousarray(a, dtype=F)
    b = np.ascontiguousarray(b, dtype=F)
    return int((a.view(np.uint32) != b.view(np.uint32)).sum())
