"""Synthetic G3M^3 workloads: nested boxes, dipole planet, solar wind.

The reference initialises from its namelist and its own planet tables
(code/3d_main.f90:651-1916); here the run is built from a small set of
physical parameters, in the reference's normalised units
(v_equiv = 1000 km/s, rho_equiv = 0.2 cm^-3 of mass rmassq,
b_equiv = v_equiv*sqrt(mu0*rho_equiv*m_p) ~ 20.5 nT, p_equiv =
rho_equiv*m_p*v_equiv^2).  Everything is host numpy, laid out as the
reference's Fortran arrays: a field ``q(nx,ny,nz,n_grids)`` is a
C-order float32 array of shape ``(n_grids, nz, ny, nx)``.

Named workloads:

* ``single61``  -- one 61x61x31 box (the reference's CPU-runnable case)
* ``nested121`` -- five nested 121x121x61 boxes, spacing 1,2,4,8,16
* ``tiny``      -- two nested 13x13x7 boxes for quick parity checks
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

F = np.float32
SPECIES = "qho"
SP_FIELDS = ("rho", "presx", "presy", "presz", "px", "py", "pz",
             "presxy", "presxz", "presyz")
#: the 34 per-cell state fields, in this package's canonical order
FIELDS = [s + f for s in SPECIES for f in SP_FIELDS] + ["epres", "bx", "by", "bz"]
NFIELDS = len(FIELDS)

V_EQUIV_KMS = 1000.0
RHO_EQUIV_CC = 0.2
B_EQUIV_NT = 20.5
P_EQUIV_PA = RHO_EQUIV_CC * 1e6 * 1.6726e-27 * (V_EQUIV_KMS * 1e3) ** 2
EV = 1.602176634e-19


@dataclass
class Workload:
    name: str
    nx: int
    ny: int
    nz: int
    ngrd: int
    steps: int
    divb_every: int
    divb_tol: float
    grid: dict                  # grd_xmin..grd_zmax, xspac: float32[ngrd]
    scal: dict                  # physical and numerical scalars
    fields: dict                # 34 state fields + bx0/by0/bz0 + rst
    lists: dict = field(default_factory=dict)   # ij*, num*, parm_* (box 1)

    @property
    def cells_per_box(self) -> int:
        return self.nx * self.ny * self.nz


def _grid(L: int, ngrd: int, wind_adjust: float = 5.0 / 4.0):
    """Box extents as code/3d_main.f90:723-736: spacing doubles per level,
    z is half as tall, the outermost box is stretched down-tail."""
    xs = np.array([2.0 ** b for b in range(ngrd)], dtype=F)
    g = {"xspac": xs}
    g["grd_xmin"] = (-L * xs).astype(F)
    g["grd_xmax"] = (L * xs).astype(F)
    g["grd_ymin"] = (-L * xs).astype(F)
    g["grd_ymax"] = (L * xs).astype(F)
    g["grd_zmin"] = (-(L / 2.0) * xs).astype(F)
    g["grd_zmax"] = ((L / 2.0) * xs).astype(F)
    diff = F(g["grd_xmax"][-1] * F(wind_adjust) - g["grd_xmax"][-1])
    g["grd_xmax"][-1] += diff
    g["grd_xmin"][-1] += diff
    return g


def _check_mating(g, ngrd):
    """Each box's lower corner must fall on a grid plane of the next
    coarser box (the reference stops otherwise, code/3d_main.f90:757-789)."""
    for b in range(ngrd - 1):
        for ax in "xyz":
            d = 1.0 + (float(g[f"grd_{ax}min"][b]) - float(g[f"grd_{ax}min"][b + 1])) \
                / float(g["xspac"][b + 1])
            if abs(d - int(d)) > 0.001:
                raise ValueError(f"box {b + 1} does not mate with box {b + 2} along {ax}")


def make_workload(name: str = "nested121", seed: int = 20261203, **over) -> Workload:
    presets = {
        "tiny": dict(L=8, ngrd=2, steps=3, divb_every=0, r_inner=1.8, eta=0.0),
        "small": dict(L=12, ngrd=2, steps=4, divb_every=0, r_inner=2.2, eta=0.0),
        "single61": dict(L=30, ngrd=1, steps=100, divb_every=10, r_inner=3.0, eta=0.0),
        "nested121": dict(L=60, ngrd=5, steps=50, divb_every=10, r_inner=3.0, eta=1e-3),
    }
    p = dict(presets[name])
    p.update(over)
    L, ngrd = int(p["L"]), int(p["ngrd"])
    nx, ny, nz = 2 * L + 1, 2 * L + 1, L + 1
    grid = _grid(L, ngrd)
    if p.get("dims") is not None:
        # a single box of any extents, centred on the planet, cell size `spacing`
        if ngrd != 1:
            raise ValueError("dims= needs a single grid")
        nx, ny, nz = (int(d) for d in p["dims"])
        h = F(p.get("spacing", 1.0))
        grid = {"xspac": np.array([h], F)}
        for ax, n in (("x", nx), ("y", ny), ("z", nz)):
            grid[f"grd_{ax}min"] = np.array([-h * F((n - 1) / 2.0)], F)
            grid[f"grd_{ax}max"] = np.array([h * F((n - 1) / 2.0)], F)
    _check_mating(grid, ngrd)
    rng = np.random.default_rng(seed)

    # --- upstream conditions (solar wind along +x in the simulation frame)
    n_sw = p.get("n_sw", 5.0)              # cm^-3
    v_sw = p.get("v_sw", 400.0)            # km/s
    t_sw = p.get("t_sw", 10.0)             # eV
    imf = p.get("imf", (0.0, 0.0, -5.0))   # nT
    srho = F(n_sw / RHO_EQUIV_CC)
    svel = F(v_sw / V_EQUIV_KMS)
    spress = F(n_sw * 1e6 * t_sw * EV / P_EQUIV_PA)
    sc = dict(
        srho=srho, spx=F(srho * svel), spy=F(0), spz=F(0), spress=spress,
        sbx_wind=F(imf[0] / B_EQUIV_NT), sby_wind=F(imf[1] / B_EQUIV_NT),
        sbz_wind=F(imf[2] / B_EQUIV_NT),
        rmassq=F(1.), rmassh=F(1.), rmasso=F(16.), gamma=F(1.6666), gamma1=F(0.6666),
        reynolds=F(256.), resist=F(64.), grav=F(p.get("grav", 0.0)), re_equiv=F(0.2),
        isotropic=False, smallbit=F(1e-5), aniso_limit=F(0.9995),
        qvlim=F(1.), hvlim=F(1.4), ovlim=F(2.), qclim=F(1.), hclim=F(1.), oclim=F(4.),
        cslim=F(1.), chirho=F(2.), chipxyz=F(2.), chierg=F(0.), difrho=F(0.008),
        diferg=F(0.004), o_conc=F(0.125), rho_frac=F(0.001), ti_te=F(1.),
        stepsz=F(0.25), t_step_max=F(20.), vlim_factor=F(0.6), mbndry=1,
        v_rot=F(0.), r_rot=F(0.), erho=F(0.), epress=F(0.), alpha_e=F(6.),
        r_inner=F(p["r_inner"]), eta=F(p["eta"]), b0=F(0.),
        # dipole tilting during the run (code/3d_main.f90:1966-1977, 2173-2187)
        tilting=bool(p.get("tilting", False)), tilt=F(p.get("tilt", 0.0)),
        dtilt=F(p.get("dtilt", 0.0)), ut=float(p.get("ut", 0.0)),
        t_equiv=F(p.get("t_equiv", 2.5)), planet_per=F(p.get("planet_per", 10.0)),
        xdip=F(0.00001), ydip=F(0.0), zdip=F(0.0),
    )
    # dipole strength: |B| = 10 b_equiv on the equator at r_inner
    b0 = F(10.0 * float(p["r_inner"]) ** 3)
    sc["b0"] = b0
    xdip = F(0.00001)

    shp = (ngrd, nz, ny, nx)
    f = {k: np.zeros(shp, F) for k in FIELDS}
    for c in ("x", "y", "z"):
        f["b" + c + "0"] = np.zeros(shp, F)
    r_in = float(p["r_inner"])
    zero_b, mid_b, srf_b = r_in - 1.5, r_in - 0.5, r_in + 0.6
    frac_h = sc["rmassh"] / sc["rmassq"]
    frac_o = sc["rmasso"] / sc["rmassq"]
    ofrac = sc["rho_frac"] * sc["o_conc"]
    for b in range(ngrd):
        dx = (grid["grd_xmax"][b] - grid["grd_xmin"][b]) / F(nx - 1)
        ax = (grid["grd_xmin"][b] + dx * np.arange(nx, dtype=F) - xdip)[None, None, :]
        ay = (grid["grd_ymin"][b] + dx * np.arange(ny, dtype=F))[None, :, None]
        az = (grid["grd_zmin"][b] + dx * np.arange(nz, dtype=F))[:, None, None]
        ar = np.sqrt(ax * ax + ay * ay + az * az).astype(F)
        ar5 = np.maximum(ar, F(1e-3)) ** 5
        bmag = -b0 / ar5
        bx0 = -3.0 * bmag * ax * az
        by0 = -3.0 * bmag * ay * az
        bz0 = bmag * (ax * ax + ay * ay - 2.0 * az * az)
        inside = ar < zero_b
        f["bx0"][b] = np.where(inside, 0, bx0)
        f["by0"][b] = np.where(inside, 0, by0)
        f["bz0"][b] = np.where(inside, 0, bz0)
        # ionospheric populations within ~5 cells of the planet
        iono = 1.0 / (1.0 + np.exp(np.minimum((ar - 5.0 * (r_in / 3.0)) / 0.7, 60.0)))
        noise = lambda: (1.0 + 0.01 * rng.uniform(-1, 1, ar.shape)).astype(F)
        rq = srho * noise()
        rh = (srho * sc["rho_frac"] * frac_h + F(1e3 / RHO_EQUIV_CC) * sc["rmassh"] * iono) * noise()
        ro = (srho * ofrac * frac_o + F(1e2 / RHO_EQUIV_CC) * sc["rmasso"] * iono) * noise()
        vx = svel * (1.0 - iono)
        pq = 0.5 * spress * np.ones_like(ar)
        ion_p = 1e6 * 1.0 * EV / P_EQUIV_PA      # 1 eV per cm^-3
        ph = 0.5 * spress * sc["rho_frac"] + ion_p * 1e3 * iono
        po = 0.5 * spress * ofrac + ion_p * 1e2 * iono
        for s, r, pr in (("q", rq, pq), ("h", rh, ph), ("o", ro, po)):
            f[s + "rho"][b] = r
            f[s + "px"][b] = r * vx
            for c in ("x", "y", "z"):
                f[s + "pres" + c][b] = pr
        f["epres"][b] = (f["qpresx"][b] + f["hpresx"][b] + f["opresx"][b]) / sc["ti_te"]

    # inner-boundary lists on box 1 (code/3d_main.f90:1383-1455 layout)
    lists = {}
    b = 0
    dx = (grid["grd_xmax"][b] - grid["grd_xmin"][b]) / F(nx - 1)
    kk, jj, ii = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ax = grid["grd_xmin"][b] + dx * ii.astype(F) - xdip
    ay = grid["grd_ymin"][b] + dx * jj.astype(F)
    az = grid["grd_zmin"][b] + dx * kk.astype(F)
    ar = np.sqrt(ax * ax + ay * ay + az * az)
    sel = {"zero": ar < zero_b, "mid": (ar >= zero_b) & (ar < mid_b),
           "srf": (ar >= mid_b) & (ar <= srf_b)}
    rst = np.zeros((1, nz, ny, nx), F)
    for nm, m in sel.items():
        k_, j_, i_ = np.nonzero(m)
        n = len(i_)
        lst = np.zeros((max(n, 1), 3, 1), np.int32)
        lst[:n, 0, 0], lst[:n, 1, 0], lst[:n, 2, 0] = i_ + 1, j_ + 1, k_ + 1
        parm = np.zeros((max(n, 1), 7, 1), F)
        for c, key in enumerate(("qrho", "hrho", "orho", "qpresx", "hpresx", "opresx", "epres")):
            parm[:n, c, 0] = f[key][0][k_, j_, i_]
        lists["ij" + nm] = lst
        lists["num" + nm] = np.array([n], np.int32)
        lists["parm_" + nm] = parm
        lists["m" + nm] = max(n, 1)
        # code/sub_set_resist.f90:31-58
        scale = {"zero": 1.0, "mid": 0.5, "srf": 0.125}[nm]
        rst[0][k_, j_, i_] = F(scale) * b0 / sc["resist"]
    f["rst"] = rst
    return Workload(name=name, nx=nx, ny=ny, nz=nz, ngrd=ngrd, steps=int(p["steps"]),
                    divb_every=int(p["divb_every"]), divb_tol=float(p.get("divb_tol", 1e-6)),
                    grid=grid, scal=sc, fields=f, lists=lists)


def precision_field(L: int = 120, seed: int = 20261205, r_inner: float = 3.0):
    """The divB precision case: one (2L+1, 2L+1, L+1) box of unit spacing
    (241x241x121 at L=120) holding a dipole field plus Gaussian noise of
    sigma = 1% of |B| on each component (seeded), as float32 arrays of shape
    (1, nz, ny, nx) -- the field whose divergence divB_correct removes."""
    nx, ny, nz = 2 * L + 1, 2 * L + 1, L + 1
    rng = np.random.default_rng(seed)
    ax = (np.arange(nx, dtype=F) - F(L) - F(0.00001))[None, None, :]
    ay = (np.arange(ny, dtype=F) - F(L))[None, :, None]
    az = (np.arange(nz, dtype=F) - F(L // 2))[:, None, None]
    ar = np.sqrt(ax * ax + ay * ay + az * az).astype(F)
    ar5 = np.maximum(ar, F(r_inner - 1.5)) ** 5
    bmag = F(-10.0 * r_inner ** 3) / ar5
    b = [(-3.0 * bmag * ax * az).astype(F), (-3.0 * bmag * ay * az).astype(F),
         (bmag * (ax * ax + ay * ay - 2.0 * az * az)).astype(F)]
    mag = np.sqrt(b[0] ** 2 + b[1] ** 2 + b[2] ** 2)
    out = []
    for c in b:
        noisy = c + F(0.01) * mag * rng.standard_normal(c.shape).astype(F)
        out.append(np.ascontiguousarray(noisy.astype(F)[None]))
    return out


def make_batch_workloads(n: int = 32, seed: int = 20261204):
    """Independent single-grid runs for batching: box sizes 41x41x21 ..
    121x121x61 (L uniform in 20..60), upstream density uniform in [1, 20]
    cm^-3, speed in [300, 800] km/s, IMF direction uniform on the sphere
    with magnitude in [2, 10] nT -- all drawn from one seeded generator."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        L = int(rng.integers(20, 61))
        n_sw = float(rng.uniform(1.0, 20.0))
        v_sw = float(rng.uniform(300.0, 800.0))
        d = rng.standard_normal(3)
        d = d / np.linalg.norm(d)
        bmag = float(rng.uniform(2.0, 10.0))
        wl = make_workload("single61", seed=seed + 1 + i, L=L, n_sw=n_sw, v_sw=v_sw,
                           imf=tuple(float(x) for x in bmag * d))
        wl.name = f"batch{i:02d}_L{L}"
        out.append(wl)
    return out
