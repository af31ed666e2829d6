"""The reference main program (code/3d_main.f90) around the native runtime.

A user of the reference keeps their ``input`` file: ``MainRun`` reads its
six namelist groups, makes the main program's set-up computations
(:720-950, :1196-1226), initialises the plasma on the device
(``g3m3_init_plasma``, :1238-1633), sizes the time step (:1861-1916) and
runs the time loop (:1925-3095) through the native runtime:

* the wall fluxes of every box at the start of each full step (:1996-2011,
  ``g3m3_flux_counter``), kept as ``fluxes.dat`` keeps them;
* the full step itself (dt sync, subcycled box-steps, final sync);
* torus injection into box 1 when ``t >= tinj`` and ``ringo`` (:3149-3295,
  ``g3m3_torus_inject``);
* a fluid (restart) file when ``t >= ts1`` (:3300-3361), followed by the
  divB clean-up when ``divb_lores`` (:3363-3408).

* the plotting tables when ``t >= tgraph`` (:3097-3144,
  ``g3m3_graphing_cuts`` and g3m3_amd.graphing, which writes the reference's
  files; the reference's Python plotting script is not started);

Spacecraft, the ``input_fluid`` upstream files and restarts from a fluid
file are not part of this loop.  The grid
shape is the reference's compile-time one (5 boxes of 121 x 121 x 61,
``limit = 60``) unless given.

Arithmetic follows the reference's declarations: ``real`` variables are
float32 (every operation rounded to single), ``real(dp)`` times are float64,
and literal constants are single precision.
"""
from __future__ import annotations

import ctypes as _ct
import os
import re
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .workload import FIELDS, Workload

F = np.float32
PI = F(np.pi)               # 4.0*atan(1.0) in single precision (code/mod_astrometry.f90:20)

NAMELISTS = ("option", "planet", "speeds", "windy", "physical", "smooth")
_INTEGERS = {"ntgraph", "ntinj"}
_STRINGS = {"run_name", "bodyname", "moonname"}
_LOGICALS = {"start", "isotropic", "write_dat", "diagnostics", "tilting", "ringo", "update",
             "reload", "divb_lores", "divb_hires", "spacecraft", "input_fluid", "warp",
             "repeat_flybys"}


def _value(name: str, text: str):
    t = text.strip()
    if name in _LOGICALS:
        v = t.lower().lstrip(".")
        if v.startswith("t"):
            return True
        if v.startswith("f"):
            return False
        raise ValueError(f"{name}: not a logical: {text}")
    if name in _STRINGS:
        return t.strip("\"'")
    if name in _INTEGERS:
        return int(t)
    return F(float(t.replace("d", "e").replace("D", "e")))


def read_input(path: str) -> dict:
    """The namelist groups option, planet, speeds, windy, physical and
    smooth of a reference ``input`` file, read in that order
    (code/3d_main.f90:703-708), in the ``$group ... $end`` form the
    reference ships.  Names left out keep the reference's static initial
    value: 0, .false. or blank."""
    text = open(path).read()
    out = {}
    pos = 0
    for grp in NAMELISTS:
        m = re.compile(r"[$&]" + grp + r"\b(.*?)[$&/]end\b", re.S | re.I).search(text, pos)
        if not m:
            raise ValueError(f"{path}: namelist group {grp} not found")
        pos = m.end()
        for name, val in re.findall(r"(\w+)\s*=\s*(\"[^\"]*\"|'[^']*'|[^\s,]+)", m.group(1)):
            name = name.lower()
            out[name] = _value(name, val)
    return out


# --- planet and moon constants (code/mod_astrometry.f90) ----------------------------
# Each entry: planet radius (km), rotation period (h), r_rot (corotation limit,
# re_equiv units), torus infall, dipole tilt (rad), dipole offset (planet
# radii) and the torus distance (the moon's orbit, planet radii).

def _saturn():
    rad = F(60268.)
    return dict(planet_rad=rad, planet_per=F(10.6562214), r_lim=F(20.0),
                torus_infall=F(0.05), planet_tilt=F(0.0), planet_xdip=F(1e-5),
                planet_ydip=F(0.0), planet_zdip=F(0.0),
                torus_dist=F(F(238.04e3) / rad))            # :114-127, :135, :236-249


def _earth():
    rad = F(6378.14)
    tilt = F(F(F(90.0) + F(80.972)) * PI / F(180.))
    return dict(planet_rad=rad, planet_per=F(24.), r_lim=F(10.0), torus_infall=F(0.),
                planet_tilt=tilt, planet_xdip=F(F(500.) / rad), planet_ydip=F(0.0),
                planet_zdip=F(0.0),
                torus_dist=F(F(384400.0) / rad))            # :160-182, :251-267


PLANETS = {"saturn": _saturn, "earth": _earth}


@dataclass
class MainSetup:
    """Everything the main program has computed when its time loop starts."""
    inp: dict
    wl: Workload
    init: dict            # the scalars of the initialisation block (g3m3_init_plasma)
    caps: tuple           # msrf, mmid, mzero
    t: float              # real(dp)
    ut: float             # real(dp)
    tmax: np.float32
    ts1: np.float32
    tgraph: np.float32
    tinj: np.float32
    deltg: np.float32
    deltinj: np.float32
    torus: dict = field(default_factory=dict)
    flux_scale: list = field(default_factory=list)
    graph: dict = field(default_factory=dict)   # write_graphing_data's scalars


def setup(inp: dict, limit: float = 60.0, n_grids: int = 5, mbndry: int = 1,
          planet: dict | None = None) -> MainSetup:
    """The main program's set-up from its input (code/3d_main.f90:720-950,
    the start branch :1196-1226), for ``n_grids`` boxes of (2 limit + 1)^2 x
    (limit + 1) cells."""
    g = {k: v for k, v in inp.items()}
    body = g.get("bodyname", "saturn") or "saturn"
    if planet is None:
        if body not in PLANETS:
            raise ValueError(f"planet {body!r}: give its constants (planet=...)")
        planet = PLANETS[body]()
    P = {k: F(v) for k, v in planet.items()}
    limit = F(limit)
    nx = ny = int(2 * limit + 1)
    nz = int(limit + 1)
    # grid (:722-749): spacing division**(i-1), the outer box stretched down-tail
    xs = np.array([2 ** i for i in range(n_grids)], F)
    grid = {"xspac": xs}
    grid["grd_xmin"] = (-(limit * xs)).astype(F)
    grid["grd_xmax"] = (limit * xs).astype(F)
    grid["grd_ymin"] = (-(limit * xs)).astype(F)
    grid["grd_ymax"] = (limit * xs).astype(F)
    grid["grd_zmin"] = (-(F(limit / F(2)) * xs)).astype(F)
    grid["grd_zmax"] = (F(limit / F(2)) * xs).astype(F)
    diff = F(F(grid["grd_xmax"][-1] * F(1.25)) - grid["grd_xmax"][-1])
    grid["grd_xmax"][-1] = F(grid["grd_xmax"][-1] + diff)
    grid["grd_xmin"][-1] = F(grid["grd_xmin"][-1] + diff)

    def v(name, default=0.0):
        x = g.get(name, default)
        return F(x) if not isinstance(x, (bool, str)) else x

    r_inner = v("r_inner")
    zero_bndry = F(r_inner + F(-1.5))
    mid_bndry = F(r_inner + F(-0.5))
    srf_bndry = F(r_inner + F(0.6))
    re_equiv, v_equiv = v("re_equiv"), v("v_equiv")
    # planet (:817-850)
    r_rot = P["r_lim"]
    v_rot = F(F(F(F(F(2.) * PI) * P["planet_rad"]) / F(P["planet_per"] * F(3600.))) / v_equiv)
    torus_inj_dist = F(P["torus_dist"] + P["torus_infall"])
    tilt = P["planet_tilt"]
    sin_tilt = F(_sinf(tilt))
    cos_tilt = F(_cosf(tilt))
    tmax_in = v("tmax")
    with np.errstate(divide="ignore", invalid="ignore"):
        dtilt = F(F(v("tilt2") - v("tilt1")) / tmax_in)
    if g.get("tilting", False):
        xdip, ydip, zdip = P["planet_xdip"], P["planet_ydip"], P["planet_zdip"]
    else:
        xdip, ydip, zdip = F(0.00001), F(0.), F(0.)
    t_equiv = F(F(P["planet_rad"] * re_equiv) / v_equiv)
    grav = F(F(v("gravity") * F(F(P["planet_rad"] * re_equiv) * F(1000.)))
             / _ipow(F(F(1000.) * v_equiv), 2))
    # list capacities (:859-861), truncated toward zero
    k = F(F(F(1.2) * F(1.33)) * PI)
    x3 = _ipow(xs[mbndry - 1], 3)
    msrf = int(F(F(k * F(_ipow(srf_bndry, 3) - _ipow(mid_bndry, 3))) / x3))
    mmid = int(F(F(k * F(_ipow(mid_bndry, 3) - _ipow(zero_bndry, 3))) / x3))
    mzero = int(F(F(k * _ipow(zero_bndry, 3)) / x3))
    # dipole strength and the upstream state (:875-943)
    rmassq, rmassh, rmasso = v("rmassq"), v("rmassh"), v("rmasso")
    erho = F(v("denh_inner") * rmassh)
    b0 = F(F(v("alf_inner1") * F(np.sqrt(erho))) * _ipow(r_inner, 3))
    gamma = v("gamma")
    gamma1 = F(gamma - F(1.))
    epress = F(F(_ipow(v("cs_inner"), 2) * erho) / gamma)
    srho = F(v("den_wind1") * rmassq)
    spress = F(F(F(_ipow(v("cs_wind"), 2) * srho) / gamma) / gamma1)
    spx = F(srho * v("vx_wind1"))
    spy = F(srho * v("vy_wind1"))
    spz = F(srho * v("vz_wind1"))
    sq = F(np.sqrt(srho))
    sbx, sby, sbz = (F(v(f"alf{c}_wind1") * sq) for c in "xyz")
    with np.errstate(divide="ignore", invalid="ignore"):
        deltg = F(tmax_in / F(g.get("ntgraph", 0)))
        deltinj = F(tmax_in / F(g.get("ntinj", 0)))
    # the start branch (:1196-1226): clock and rotation
    if not g.get("start", False):
        raise ValueError("restarts from a fluid file (start = .false.) are not supported here")
    ut = float(v("utstart"))
    t = 0.0
    if v("utstart") > F(0.00001):
        t = ut * 3600.0 / float(t_equiv)
    ts1 = F(t + float(v("tsave")))
    tmax = F(t + float(tmax_in))
    tgraph = F(t + float(deltg))
    tinj = F(t + float(deltinj))
    per = P["planet_per"]
    rot_angle = F(0.)
    if per != 0:
        nrot = int(ut / float(per))
        rot_hrs = F(ut - float(F(per * F(nrot))))
        rot_angle = F(F(F(F(2.) * PI) * rot_hrs) / per)
    init = dict(rmassq=rmassq, rmassh=rmassh, rmasso=rmasso, ti_te=v("ti_te"),
                re_equiv=re_equiv, r_inner=r_inner, zero_bndry=zero_bndry, mid_bndry=mid_bndry,
                srf_bndry=srf_bndry, alpha_e=v("alpha_e"), erho=erho, d_min=F(0.01),
                cs_inner=v("cs_inner"), srho=srho, spx=spx, spy=spy, spz=spz, spress=spress,
                cs_wind=v("cs_wind"), gamma=gamma, rho_frac=v("rho_frac"), resist=v("resist"),
                v_rot=v_rot, r_rot=r_rot, rot_angle=rot_angle, xdip=xdip, ydip=ydip,
                zdip=zdip, sin_tilt=sin_tilt, cos_tilt=cos_tilt, b0=b0)
    scal = dict(
        rmassq=rmassq, rmassh=rmassh, rmasso=rmasso, gamma=gamma, gamma1=gamma1,
        reynolds=v("reynolds"), resist=v("resist"), grav=grav, re_equiv=re_equiv,
        isotropic=bool(g.get("isotropic", False)), smallbit=F(1e-5),
        aniso_limit=v("aniso_limit"), qvlim=v("qvlim"), hvlim=v("hvlim"), ovlim=v("ovlim"),
        qclim=v("qclim"), hclim=v("hclim"), oclim=v("oclim"), cslim=v("cslim"),
        chirho=v("chirho"), chipxyz=v("chipxyz"), chierg=v("chierg"), difrho=v("difrho"),
        diferg=v("diferg"), o_conc=v("o_conc"), rho_frac=v("rho_frac"), ti_te=v("ti_te"),
        stepsz=v("stepsz"), t_step_max=F(20.0), vlim_factor=F(0.6), mbndry=mbndry,
        srho=srho, spx=spx, spy=spy, spz=spz, spress=spress, sbx_wind=sbx, sby_wind=sby,
        sbz_wind=sbz, v_rot=v_rot, r_rot=r_rot, erho=erho, epress=epress, alpha_e=v("alpha_e"),
        r_inner=r_inner, eta=F(0.), b0=b0, tilting=bool(g.get("tilting", False)), tilt=tilt,
        dtilt=dtilt, ut=ut, t_equiv=t_equiv, planet_per=per, xdip=xdip, ydip=ydip, zdip=zdip)
    shp = (n_grids, nz, ny, nx)
    fields = {k: np.zeros(shp, F) for k in FIELDS}
    for c in "xyz":
        fields[f"b{c}0"] = np.zeros(shp, F)
    fields["rst"] = np.zeros((mbndry, nz, ny, nx), F)
    lists = dict(msrf=msrf, mmid=mmid, mzero=mzero)
    for nm, m in (("srf", msrf), ("mid", mmid), ("zero", mzero)):
        lists["ij" + nm] = np.zeros((m, 3, mbndry), np.int32)
        lists["parm_" + nm] = np.zeros((m, 7, mbndry), F)
        lists["num" + nm] = np.zeros(mbndry, np.int32)
    wl = Workload(name=f"main:{body}", nx=nx, ny=ny, nz=nz, ngrd=n_grids, steps=0,
                  divb_every=0, divb_tol=0.0, grid=grid, scal=scal, fields=fields, lists=lists)
    torus = dict(xdip=xdip, ydip=ydip, zdip=zdip, re_equiv=re_equiv, v_rot=v_rot,
                 torus_inj_dist=torus_inj_dist, torus_rad=v("torus_rad"), srf_bndry=srf_bndry,
                 denq_torus=v("denq_torus"), denh_torus=v("denh_torus"),
                 deno_torus=v("deno_torus"), rmassq=rmassq, rmassh=rmassh, rmasso=rmasso,
                 t_torus=v("t_torus"), aniso_factor=v("aniso_factor"), reduct=v("reduct"))
    # wall-flux scale per box (:1998-1999)
    rho_equiv = v("rho_equiv")
    scale = [F(F(F(rho_equiv * v_equiv) * F(1.e5))
               * _ipow(F(F(F(xs[b] * P["planet_rad"]) * re_equiv) * F(1.e5)), 2))
             for b in range(n_grids)]
    # b_equiv in nT (:907): v_equiv*1.e3 * sqrt(mu0 * rho_equiv*1.e6*m_prot) * 1.e9
    mu0, m_prot = F(F(4.e-7) * PI), F(1.67262e-27)
    b_equiv = F(F(F(v_equiv * F(1.e3)) * F(np.sqrt(F(F(F(mu0 * rho_equiv) * F(1.e6)) * m_prot))))
                * F(1.e9))
    graph = dict(run_name=str(g.get("run_name", "")), limit=limit, re_equiv=re_equiv,
                 b_equiv=b_equiv, v_equiv=v_equiv, rho_equiv=rho_equiv,
                 planet_rad=P["planet_rad"])
    return MainSetup(inp=inp, wl=wl, init=init, caps=(msrf, mmid, mzero), t=t, ut=ut,
                     tmax=tmax, ts1=ts1, tgraph=tgraph, tinj=tinj, deltg=deltg,
                     deltinj=deltinj, torus=torus, flux_scale=scale, graph=graph)


_LIBM = _ct.CDLL("libm.so.6")
_LIBM.sinf.restype = _LIBM.cosf.restype = _ct.c_float
_LIBM.sinf.argtypes = _LIBM.cosf.argtypes = [_ct.c_float]


def _sinf(x):
    return F(_LIBM.sinf(float(F(x))))


def _cosf(x):
    return F(_LIBM.cosf(float(F(x))))


def _ipow(x, n: int):
    """x**n for an integer n as the compiled reference evaluates it: left to
    right products in single precision."""
    x = F(x)
    r = x
    for _ in range(n - 1):
        r = F(r * x)
    return r


_INIT_ORDER = ("rmassq", "rmassh", "rmasso", "ti_te", "re_equiv", "r_inner", "zero_bndry",
               "mid_bndry", "srf_bndry", "alpha_e", "erho", "d_min", "cs_inner", "srho", "spx",
               "spy", "spz", "spress", "cs_wind", "gamma", "rho_frac", "resist", "v_rot",
               "r_rot", "rot_angle", "xdip", "ydip", "zdip", "sin_tilt", "cos_tilt", "b0")
_TORUS_ORDER = ("xdip", "ydip", "zdip", "re_equiv", "v_rot", "torus_inj_dist", "torus_rad",
                "srf_bndry", "denq_torus", "denh_torus", "deno_torus", "rmassq", "rmassh",
                "rmasso", "t_torus", "aniso_factor", "reduct")


def init_on_device(S: MainSetup) -> None:
    """The initialisation block (code/3d_main.f90:1238-1633) on the device:
    fills S.wl's fields, b0, resistivity and boundary lists in place."""
    import torch
    wl = S.wl
    msrf, mmid, mzero = S.caps
    dev = {k: torch.zeros(v.shape, dtype=torch.float32, device="cuda")
           for k, v in wl.fields.items()}
    lk = ("parm_srf", "parm_mid", "parm_zero", "ijsrf", "ijmid", "ijzero")
    dl = {k: torch.from_numpy(wl.lists[k]).cuda() for k in lk}
    num = {k: wl.lists[k] for k in ("numsrf", "nummid", "numzero")}
    g = wl.grid
    gmin = np.ascontiguousarray(np.stack([g["grd_xmin"], g["grd_ymin"], g["grd_zmin"]], 1), F)
    gmax = np.ascontiguousarray(np.stack([g["grd_xmax"], g["grd_ymax"], g["grd_zmax"]], 1), F)
    _lib.call("init_plasma", *[dev[k] for k in FIELDS], dev["bx0"], dev["by0"], dev["bz0"],
              dev["rst"], wl.nx, wl.ny, wl.nz, wl.ngrd, int(wl.scal["mbndry"]), msrf, mmid,
              mzero, gmin, gmax, dl["parm_srf"], dl["parm_mid"], dl["parm_zero"], dl["ijsrf"],
              dl["ijmid"], dl["ijzero"], num["numsrf"], num["nummid"], num["numzero"],
              *[F(S.init[k]) for k in _INIT_ORDER])
    _lib.sync()
    for k, t in dev.items():
        wl.fields[k] = t.cpu().numpy()
    for k, t in dl.items():
        wl.lists[k] = t.cpu().numpy()


class MainRun:
    """The reference's main program for a run from scratch, advanced by the
    native runtime.  ``run_loop()`` executes the time loop to tmax (or
    ``max_steps`` full steps); fluid files go to ``out_dir`` as fluidNN.
    Logs, as the reference writes them:

    * ``step_log``: (box, t, t_old, t_new) per box-step (speeds.dat);
    * ``flux_log``: (ut, box, q in, h in, o in, q out, h out, o out) per box
      and full step (fluxes.dat);
    * ``torus_log``: (t, q, h, o) injected totals per injection."""

    def __init__(self, input_path: str, out_dir: str = ".", **setup_kw):
        from .driver import FastRun
        self.S = setup(read_input(input_path), **setup_kw)
        init_on_device(self.S)
        self.run = FastRun(self.S.wl)
        self.run._L.g3m3_run_set_time(_ct.c_void_p(self.run.h), _ct.c_double(self.S.t))
        self.run.t = self.S.t
        self.run.init_dt()
        self.out_dir = out_dir
        self.fluid_f, self.offset_f = 11, -10        # :296-297, :1209
        self.ts1, self.tmax = self.S.ts1, self.S.tmax
        self.tgraph, self.tinj = self.S.tgraph, self.S.tinj
        self.step_log, self.flux_log, self.torus_log, self.fluid_files = [], [], [], []
        self.graph_log = []
        self.steps = 0

    def _grid_args(self):
        r = self.run
        return [r.a[k] for k in ("grd_xmin", "grd_xmax", "grd_ymin", "grd_ymax", "grd_zmin",
                                 "grd_zmax")]

    def _wall_fluxes(self):
        """flux_counter of every box, n_grids down to 1 (:1997-2011)."""
        r = self.run
        r.sync_arrays()
        rows = []
        for box in range(r.ngrd, 0, -1):
            out = [np.zeros(1, F) for _ in range(6)]
            _lib.call("flux_counter", *[r.a[k] for k in ("qpx", "qpy", "qpz", "hpx", "hpy",
                                                           "hpz", "opx", "opy", "opz")],
                      r.nx, r.ny, r.nz, r.ngrd, box, F(r.s["rmassq"]), F(r.s["rmassh"]),
                      F(r.s["rmasso"]), self.S.flux_scale[box - 1], *out)
            qi, qo, hi, ho, oi, oo = (float(o[0]) for o in out)
            rows.append((box, qi, hi, oi, qo, ho, oo))
        _lib.check("flux_counter")
        return rows

    def _torus(self):
        """code/3d_main.f90:3149-3295, into box 1."""
        r = self.run
        r.sync_arrays()
        T = self.S.torus
        tot = [np.zeros(1, F) for _ in range(3)]
        g = r.wl.grid
        gmin = np.ascontiguousarray(np.stack([g["grd_xmin"], g["grd_ymin"], g["grd_zmin"]], 1),
                                    F)
        gmax = np.ascontiguousarray(np.stack([g["grd_xmax"], g["grd_ymax"], g["grd_zmax"]], 1),
                                    F)
        _lib.call("torus_inject", *[r.a[k] for k in FIELDS], r.a["bx0"], r.a["by0"],
                  r.a["bz0"], r.nx, r.ny, r.nz, r.ngrd, gmin, gmax,
                  *[F(T[k]) for k in _TORUS_ORDER], *tot)
        _lib.sync()
        self.torus_log.append((r.t, *[float(x[0]) for x in tot]))

    def _graph(self, ut):
        """code/3d_main.f90:3097-3144: the plotting tables (write_graphing_data)
        under out_dir/figures/data."""
        from .graphing import write_graphing_data
        G = self.S.graph
        n = write_graphing_data(self.run, self.out_dir, G["run_name"], ut, G["limit"],
                                G["re_equiv"], G["b_equiv"], G["v_equiv"], G["rho_equiv"],
                                G["planet_rad"])
        self.graph_log.append((self.run.t, n))

    def _write_fluid(self):
        """code/3d_main.f90:3304-3361."""
        path = os.path.join(self.out_dir, f"fluid{self.fluid_f + self.offset_f:02d}")
        self.run.write_fluid(path)
        self.fluid_files.append(path)
        self.fluid_f += 1
        if self.fluid_f > 20:
            self.fluid_f = 11
            self.offset_f += 10
            if self.offset_f > 89:
                self.tmax = F(self.run.t - 1.0)

    def _divb(self):
        """code/3d_main.f90:3363-3408: divB_correct twice per box (the
        reference's solver and tolerance), then flanks_synced on b and corer."""
        from . import signatures as sig
        r = self.run
        for box in range(r.ngrd, 0, -1):
            for _ in range(2):
                r.divb(box, F(1.0e-4), 0)
        r.sync_arrays()
        for box in range(r.ngrd - 1, 0, -1):
            for c in ("bx", "by", "bz"):
                _lib.call("flanks_synced", r.a[c], r.nx, r.ny, r.nz, r.ngrd, box,
                          *self._grid_args())
        for box in range(1, r.ngrd):
            _lib.call("bndry_corer_v", r.ptrs("", sig.CORER_ORDER), r.nx, r.ny, r.nz, r.ngrd,
                      box, *self._grid_args())
        _lib.sync()

    # -- the time loop (code/3d_main.f90:1925-3095) -----------------------------------
    def run_loop(self, max_steps: int | None = None):
        r = self.run
        t_equiv = float(F(self.S.wl.scal["t_equiv"]))
        while r.t < float(self.tmax):
            if max_steps is not None and self.steps >= max_steps:
                break
            # the fluxes are counted after the clock update, before any
            # box-step: the state is the previous step's, ut the new one
            fluxes = self._wall_fluxes()
            r.full_step()
            self.steps += 1
            ut = r.t * t_equiv / 3600.0
            self.flux_log += [(ut, *row) for row in fluxes]
            self.step_log += [(box, r.t, t_old, t_new) for box, t_old, t_new in r.step_log()]
            if r.t >= float(self.tgraph):
                self._graph(ut)
                self.tgraph = F(self.tgraph + self.S.deltg)
            if r.t >= float(self.tinj) and self.S.inp.get("ringo", False):
                self._torus()
                self.tinj = F(self.tinj + self.S.deltinj)
            if r.t >= float(self.ts1):
                self._write_fluid()
                self.ts1 = F(self.ts1 + F(self.S.inp.get("tsave", 0.0)))
                if self.S.inp.get("divb_lores", False):
                    self._divb()
        r.sync_arrays()
        return self


def main(argv=None):
    """python -m g3m3_amd.mainprog INPUT [OUT_DIR] [--steps N]"""
    import argparse
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("input")
    ap.add_argument("out_dir", nargs="?", default=".")
    ap.add_argument("--steps", type=int, default=None)
    a = ap.parse_args(argv)
    m = MainRun(a.input, a.out_dir).run_loop(a.steps)
    print(f"{m.steps} full steps, t = {m.run.t!r}, {m.run.box_steps} box-steps, "
          f"fluid files: {', '.join(os.path.basename(f) for f in m.fluid_files) or 'none'}")


if __name__ == "__main__":
    main()
