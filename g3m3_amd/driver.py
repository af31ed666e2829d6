"""Host driver of the MI355X hot path: the reference's box-step and
full-step sequence (code/3d_main.f90:1925-3095) over the C-ABI.

All per-cell arrays live in HBM for the whole run (torch tensors used as
plain device allocations); only the small arrays the library reads on the
host (grid extents, xspac, list counts, t_old/t_new) are numpy.  The state,
``old`` and ``wrk`` sets are each one contiguous [34, n_grids, nz, ny, nx]
allocation; a field is a contiguous view of it, laid out as the
reference's Fortran ``q(nx,ny,nz,n_grids)``.

One device->host transfer per full step: the CFL maxima of each box, read
after its last validation pass (the reference reads them into
``t_stepnew``, code/3d_main.f90:3038).
"""
from __future__ import annotations

import ctypes as _ct
import ctypes.util as _ctu

import numpy as np

from . import _lib
from . import signatures as sig
from .workload import FIELDS, Workload

F = np.float32
SCRATCH3 = ("curx", "cury", "curz", "evx", "evy", "evz", "tvx", "tvy", "tvz",
            "efldx", "efldy", "efldz", "bsx", "bsy", "bsz", "btot", "vvx", "vvy", "vvz")


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise _lib.G3M3Error("no GPU visible: the G3M3 hot path runs only on the device")
    return torch


# the C library's single-precision sin / cos, as a compiled driver calls them
_LIBM = _ct.CDLL(_ctu.find_library("m"))
for _fn in ("sinf", "cosf"):
    getattr(_LIBM, _fn).restype = _ct.c_float
    getattr(_LIBM, _fn).argtypes = [_ct.c_float]


class Run:
    """One nested-box run resident on the GPU."""

    def __init__(self, wl: Workload):
        torch = _torch()
        self.torch = torch
        self.wl = wl
        self.nx, self.ny, self.nz, self.ngrd = wl.nx, wl.ny, wl.nz, wl.ngrd
        dev = torch.device("cuda")
        shp = (len(FIELDS), self.ngrd, self.nz, self.ny, self.nx)
        self.sets = {}
        a = {}
        for pre in ("", "old", "wrk"):
            t = torch.zeros(shp, dtype=torch.float32, device=dev)
            self.sets[pre] = t
            for f, k in enumerate(FIELDS):
                a[pre + k] = t[f]
        host = np.stack([wl.fields[k] for k in FIELDS])
        self.sets[""].copy_(torch.from_numpy(host))
        for k in ("bx0", "by0", "bz0", "rst"):
            a[k] = torch.from_numpy(np.ascontiguousarray(wl.fields[k])).to(dev)
        for k in SCRATCH3:
            a[k] = torch.zeros((self.nz, self.ny, self.nx), dtype=torch.float32, device=dev)
        for k, v in wl.grid.items():
            a[k] = np.ascontiguousarray(v, dtype=F)
        for k in ("ijsrf", "ijmid", "ijzero", "parm_srf", "parm_mid", "parm_zero"):
            a[k] = torch.from_numpy(np.ascontiguousarray(wl.lists[k])).to(dev)
        for k in ("numsrf", "nummid", "numzero"):
            a[k] = np.ascontiguousarray(wl.lists[k], dtype=np.int32)
        self.t_old = np.zeros(self.ngrd, np.float64)
        self.t_new = np.zeros(self.ngrd, np.float64)
        a["t_old"], a["t_new"] = self.t_old, self.t_new
        for k in ("pxmax", "pymax", "pzmax", "csmax", "alfmax", "fastest"):
            a[k] = np.zeros(1, F)
        self.a = a
        s = dict(wl.scal)
        s.update(nx=self.nx, ny=self.ny, nz=self.nz, n_grids=self.ngrd,
                 msrf=int(wl.lists["msrf"]), mmid=int(wl.lists["mmid"]),
                 mzero=int(wl.lists["mzero"]))
        self.s = s
        self.t_step = np.zeros(self.ngrd, np.float64)
        self.t_stepnew = np.zeros(self.ngrd, np.float64)
        self.t = 0.0
        # the main loop's tilt and universal time (code/3d_main.f90:1966-1977)
        self.tilt = self.old_tilt = F(s.get("tilt", 0.0))
        self.ut = self.utold = float(s.get("ut", 0.0))
        self.cell_updates = 0
        self.box_steps = 0
        # host arrays of device pointers for the 34-field entry points
        self._ptr_cache = {}
        _lib.lib()

    # -- plumbing -------------------------------------------------------------
    def args(self, names):
        out = []
        for nm in names:
            if isinstance(nm, str):
                out.append(self.a[nm] if nm in self.a else self.s[nm])
            else:
                out.append(nm)
        return out

    def call(self, name, names):
        _lib.call(name, *self.args(names))

    def ptrs(self, pre, order):
        key = (pre, tuple(order))
        p = self._ptr_cache.get(key)
        if p is None:
            p = np.array([self.a[pre + k].data_ptr() for k in order], dtype=np.uint64)
            self._ptr_cache[key] = p
        return p

    # -- pieces ------------------------------------------------------------------
    def _efield(self, pre):
        self.call("calcur", [pre + "bx", pre + "by", pre + "bz", "nx", "ny", "nz", "n_grids",
                             "box", "curx", "cury", "curz", "rx", "ry", "rz"])
        self.call("bfield_total", [pre + "bx", pre + "by", pre + "bz", "bx0", "by0", "bz0",
                                   "bsx", "bsy", "bsz", "btot", "nx", "ny", "nz", "n_grids",
                                   "box", "smallbit"])
        self.call("fnd_evel", sig.fnd_evel(pre))
        if float(self.s["eta"]) != 0.0:
            self.call("bande_eta", sig.bande(pre, eta=True))
        else:
            self.call("bande", sig.bande(pre))

    def _push(self, out, old, src, delt):
        self.s["delt"] = F(delt)
        self.call("push_elec", ["wrkepres", "oldepres", "epres", "evx", "evy", "evz", "gamma",
                                "gamma1", "nx", "ny", "nz", "n_grids", "box", "delt",
                                "rx", "ry", "rz"])
        for sp, rm in zip("qho", ("rmassq", "rmassh", "rmasso")):
            self.call("push_ion", sig.push_ion(out, old, src, sp, rm))
        self.call("push_bfld", [out + "bx", out + "by", out + "bz", old + "bx", old + "by",
                                old + "bz", "efldx", "efldy", "efldz", "nx", "ny", "nz",
                                "n_grids", "box", "delt", "rx", "ry", "rz"])

    def _bounds(self, pre, box, tgt, src, t_grid=None):
        if box == self.ngrd:
            self.call("bndry_outer", sig.bndry_outer(pre))
        else:
            _lib.call("bndry_flanks_v", self.ptrs(tgt, sig.FLANK_ORDER),
                      self.ptrs(src, sig.FLANK_ORDER), self.ptrs("old", sig.FLANK_ORDER),
                      self.a["vvx"], self.nx, self.ny, self.nz, self.ngrd, box,
                      self.t_old, self.t_new, _lib.F64(t_grid), *[self.a[g] for g in sig.GRID])
        if box <= self.s["mbndry"]:
            self.call("bndry_inner", sig.bndry_inner(pre))

    def _validate(self, pre, box):
        self.call("set_rho", sig.set_rho(pre))
        self.s["vlim"] = F(self.s["vlim_factor"]) * F(box)
        self.call("set_speed_agrd", sig.set_speed(pre))

    def _store_old(self, b):
        # store_array x34 (code/3d_main.f90:2193-2229): old(:,:,:,box) =
        # state(:,:,:,box), field by field, through the library's entry point
        for k in FIELDS:
            self.call("store_array", ["old" + k, k, "nx", "ny", "nz", "n_grids", "box"])

    # -- code/3d_main.f90:2163-3065 ---------------------------------------------------
    def box_step(self, box):
        a, s = self.a, self.s
        b = box - 1
        self.t_old[b] = self.t_new[b]
        self.t_new[b] = self.t_old[b] + self.t_step[b]
        delt = F(self.t_step[b])
        delt2 = delt / F(2.)
        s["box"] = box
        if s.get("tilting"):
            self._tilt(box, delt2)
        self._store_old(b)
        s["rx"] = s["ry"] = s["rz"] = F(a["xspac"][b])

        def clamp(tg):
            if tg > self.t_new[b + 1]:
                tg = self.t_new[b + 1]
            if tg < self.t_old[b + 1]:
                tg = self.t_old[b + 1]
            return tg

        # step 1
        self._efield("")
        self._push("wrk", "old", "", F(0.5) * delt)
        tg = clamp(self.t_old[b] + float(delt2)) if box < self.ngrd else None
        self._bounds("wrk", box, "wrk", "", tg)
        self._validate("wrk", box)
        # step 2
        if s.get("tilting"):
            self._tilt(box, delt)
        self._efield("wrk")
        self._push("", "old", "wrk", delt)
        tg = clamp(self.t_old[b] + float(delt)) if box < self.ngrd else None
        self._bounds("", box, "wrk", "", tg)
        self._validate("", box)
        # Lapidus
        s["delt"] = delt
        for sp in "qho":
            self.call("species_velocity", [sp + "rho", sp + "px", sp + "py", sp + "pz",
                                           "curx", "cury", "curz", "nx", "ny", "nz",
                                           "n_grids", "box", "smallbit"])
            self.call("lap_plasma", sig.lap_plasma(sp, "", "wrk", ("curx", "cury", "curz")))
        self.call("fnd_vtot", ["qpx", "qpy", "qpz", "qrho", "hpx", "hpy", "hpz", "hrho",
                               "opx", "opy", "opz", "orho", "curx", "cury", "curz",
                               "nx", "ny", "nz", "n_grids", "box", "rmassq", "rmassh", "rmasso"])
        self.call("lap_elec", ["epres", "wrkepres", "curx", "cury", "curz", "nx", "ny", "nz",
                               "n_grids", "box", "chirho", "chipxyz", "chierg", "delt"])
        self.call("lap_bfld", ["bx", "by", "bz", "wrkbx", "wrkby", "wrkbz", "curx", "cury",
                               "curz", "nx", "ny", "nz", "n_grids", "box", "chirho",
                               "chipxyz", "chierg", "delt"])
        self._bounds("wrk", box, "wrk", "", tg)
        self._validate("wrk", box)
        # FCT
        self.call("flux_correct", sig.flux_correct())
        self._bounds("", box, "", "", tg)
        self._validate("", box)
        fastest = a["fastest"][0]
        self.t_stepnew[b] = min(F(s["stepsz"]) * F(a["xspac"][b]) / fastest, F(s["t_step_max"]))
        if box < self.ngrd:
            if abs(self.t_new[b] - self.t_new[b + 1]) <= 0.005 * self.t_step[b]:
                _lib.call("bndry_corer_v", self.ptrs("", sig.CORER_ORDER), self.nx, self.ny,
                          self.nz, self.ngrd, box, *[a[g] for g in sig.GRID])
                self.t_old[b + 1] = self.t_new[b + 1]
        self.cell_updates += self.nx * self.ny * self.nz
        self.box_steps += 1

    # -- code/3d_main.f90:1861-1903 --------------------------------------------------
    def init_dt(self):
        s = self.s
        for box in range(self.ngrd, 0, -1):
            b = box - 1
            self.t_old[b] = 0.
            self.t_new[b] = 0.
            s["box"] = box
            self._validate("", box)
            self.t_stepnew[b] = min(F(s["stepsz"]) * F(self.a["xspac"][b])
                                    / self.a["fastest"][0], F(s["t_step_max"]))
        _lib.check("init_dt")

    # -- code/3d_main.f90:1925-3095 ----------------------------------------------------
    def full_step(self, dt_override=None):
        n = self.ngrd
        ts = self.t_step
        smallest = F(0)
        m_small = n
        for box in range(n, 0, -1):
            b = box - 1
            ts[b] = self.t_stepnew[b]
            self.t_old[b] = 0.
            self.t_new[b] = 0.
            if box == n:
                smallest = F(ts[b])
                m_small = box
            else:
                if ts[b] > ts[b + 1]:
                    ts[b] = ts[b + 1]
                if float(smallest) > ts[b]:
                    smallest = F(ts[b])
                    m_small = box
        for box in range(1, n + 1):
            b = box - 1
            if box <= m_small:
                ts[b] = float(smallest)
            else:
                astep = F(ts[b] / ts[b - 1] + 0.50)
                nsteps = min(max(int(astep), 1), 2)
                ts[b] = ts[b - 1] * nsteps
        if dt_override is not None:
            ts[:] = dt_override
        # the main loop stops on a step below t_step_min = 1e-6
        # (code/3d_main.f90:116, 1969-1973); a step that is not finite too
        dm = F(ts[n - 1])
        if not (np.isfinite(dm) and dm >= F(1.e-6)) or not (np.isfinite(ts[m_small - 1])
                                                             and ts[m_small - 1] > 0):
            raise _lib.G3M3Error(f"time step {float(dm)} below t_step_min 1e-6 or not finite")
        m_step = int(ts[n - 1] / ts[m_small - 1])
        # the main loop's clock (code/3d_main.f90:1966-1977): delt single,
        # t and ut double, tilt single
        delt = F(ts[n - 1])
        self.utold, self.old_tilt = self.ut, self.tilt
        self.t += float(delt)
        self.ut = self.t * float(F(self.s.get("t_equiv", 1.0))) / 3600.0
        self.tilt = F(self.tilt + F(F(self.s.get("dtilt", 0.0)) * delt))
        for _ in range(m_step):
            for box in range(n, 0, -1):
                b = box - 1
                yes = box == 1 or (self.t_old[b] == self.t_new[b]
                                   and self.t_new[b] < ts[n - 1]
                                   and abs(self.t_new[b] - self.t_new[0]) <= 0.005 * ts[0])
                if yes:
                    self.box_step(box)
        self.t_old[0] = self.t_new[0]
        for box in range(1, n):
            self.s["box"] = box
            _lib.call("bndry_corer_v", self.ptrs("", sig.CORER_ORDER), self.nx, self.ny,
                      self.nz, self.ngrd, box, *[self.a[g] for g in sig.GRID])
            self.call("set_rho", sig.set_rho(""))
        _lib.check("full_step")

    def _tilt(self, box, dt_off):
        """The box's b0 recomputed for the tilt and the planet's rotation at
        t_old + dt_off: the half step before step 1 (code/3d_main.f90:
        2173-2187), the full step before step 2 (:2465-2478).  t_old and
        utold are double, the rest single; sin / cos are the C library's
        sinf / cosf."""
        s, a = self.s, self.a
        pi = F(np.pi)   # 4*atan(1.0) in single precision
        tb = float(self.t_old[box - 1]) + float(F(dt_off))
        atilt = F(float(self.old_tilt) + tb * float(F(s["dtilt"])))
        arg = F(F(atilt * pi) / F(180.0))
        sin_tilt, cos_tilt = _LIBM.sinf(float(arg)), _LIBM.cosf(float(arg))
        ut2 = F(self.utold + tb * float(F(s["t_equiv"])) / 3600.0)
        per = F(s["planet_per"])
        rot_angle = F(0.0)   # a planet without rotation (period 0)
        if per != 0:
            nrot2 = int(F(ut2 / per))
            rot_hr2 = F(ut2 - F(F(nrot2) * per))
            rot_angle = F(F(F(F(2.0) * pi) * rot_hr2) / per)
        _lib.call("mak_dip_grd", a["bx0"], a["by0"], a["bz0"], self.nx, self.ny, self.nz,
                  self.ngrd, int(s["mbndry"]), box, a["ijzero"], a["numzero"],
                  int(s["mzero"]), F(s["r_inner"]), *[a[g] for g in sig.GRID],
                  rot_angle, F(s["xdip"]), F(s["ydip"]), F(s["zdip"]), F(sin_tilt),
                  F(cos_tilt), F(s["b0"]))

    def divb(self, box, errlim, mode=0, max_iters=0):
        """divB_correct on one box (code/sub_divB_correct.f90:25-175)."""
        import ctypes
        it = ctypes.c_longlong(0)
        r2 = ctypes.c_double(0.0)
        st = _lib.call("divb_correct_ex", self.nx, self.ny, self.nz, self.ngrd, box,
                       self.a["xspac"], F(errlim), self.a["bx"], self.a["by"], self.a["bz"],
                       int(mode), ctypes.c_longlong(max_iters), it, r2)
        _lib.check("divb")
        return st, it.value, r2.value

    def state_host(self, prefix=""):
        """The 34 fields (or their old / wrk arrays: prefix "old", "wrk")."""
        return {k: self.a[prefix + k].cpu().numpy() for k in FIELDS}


# -----------------------------------------------------------------------------
# Native runtime (g3m3_run_*): the same full step, issued from C++ with the
# fused kernels; Python makes one call per full step.


_MAXG = 8
_SCALARS = ("rmassq", "rmassh", "rmasso", "gamma", "gamma1", "reynolds", "resist", "grav",
            "re_equiv", "smallbit", "aniso_limit", "qvlim", "hvlim", "ovlim", "qclim", "hclim",
            "oclim", "cslim", "chirho", "chipxyz", "chierg", "difrho", "diferg", "o_conc",
            "rho_frac", "ti_te", "stepsz", "t_step_max", "vlim_factor", "srho", "spx", "spy",
            "spz", "spress", "sbx_wind", "sby_wind", "sbz_wind", "v_rot", "r_rot", "eta")


class RunDesc(_ct.Structure):
    """Mirror of g3m3_run_desc (include/g3m3.h)."""
    _fields_ = ([("nx", _ct.c_int), ("ny", _ct.c_int), ("nz", _ct.c_int), ("n_grids", _ct.c_int),
                 ("state", _ct.c_void_p * 34), ("old", _ct.c_void_p * 34),
                 ("wrk", _ct.c_void_p * 34)]
                + [(k, _ct.c_void_p) for k in ("bx0", "by0", "bz0", "rst", "ijsrf", "ijmid",
                                               "ijzero", "parm_srf", "parm_mid", "parm_zero")]
                + [(k, _ct.c_int) for k in ("numsrf", "nummid", "numzero", "mbndry")]
                + [(k, _ct.c_float * _MAXG) for k in ("grd_xmin", "grd_xmax", "grd_ymin",
                                                      "grd_ymax", "grd_zmin", "grd_zmax",
                                                      "xspac")]
                + [(k, _ct.c_float) for k in _SCALARS]
                + [("isotropic", _ct.c_int), ("tilting", _ct.c_int), ("tilt", _ct.c_float),
                   ("dtilt", _ct.c_float), ("ut", _ct.c_double), ("t_equiv", _ct.c_float),
                   ("planet_per", _ct.c_float), ("xdip", _ct.c_float), ("ydip", _ct.c_float),
                   ("zdip", _ct.c_float), ("dip_b0", _ct.c_float), ("r_inner", _ct.c_float)])


# the reference's fluid-file array order (code/3d_main.f90:3312-3348)
FLUID_ORDER = [sp + f for sp in "qho" for f in ("rho", "px", "py", "pz", "presx", "presy",
                                                   "presz", "presxy", "presxz", "presyz")] + \
    ["bx", "by", "bz", "epres", "bx0", "by0", "bz0"]
_LIST_KEYS = ["parm_srf", "parm_mid", "parm_zero", "ijsrf", "ijmid", "ijzero", "numsrf",
              "nummid", "numzero"]


class FastRun(Run):
    """A run advanced by the native runtime (one C-ABI call per full step)."""

    def __init__(self, wl: Workload):
        super().__init__(wl)
        L = _lib.lib()
        L.g3m3_run_create.restype = _ct.c_void_p
        L.g3m3_run_create.argtypes = [_ct.POINTER(RunDesc)]
        for nm in ("g3m3_run_init_dt", "g3m3_run_full_step", "g3m3_run_box_step",
                   "g3m3_run_sync_arrays", "g3m3_run_full_step_async", "g3m3_run_finish"):
            getattr(L, nm).restype = _ct.c_int
        d = RunDesc()
        d.nx, d.ny, d.nz, d.n_grids = self.nx, self.ny, self.nz, self.ngrd
        for f, k in enumerate(FIELDS):
            d.state[f] = self.a[k].data_ptr()
            d.old[f] = self.a["old" + k].data_ptr()
            d.wrk[f] = self.a["wrk" + k].data_ptr()
        for k in ("bx0", "by0", "bz0", "rst", "ijsrf", "ijmid", "ijzero", "parm_srf",
                  "parm_mid", "parm_zero"):
            setattr(d, k, self.a[k].data_ptr())
        d.numsrf = int(self.a["numsrf"][0])
        d.nummid = int(self.a["nummid"][0])
        d.numzero = int(self.a["numzero"][0])
        d.mbndry = int(self.s["mbndry"])
        for k in ("grd_xmin", "grd_xmax", "grd_ymin", "grd_ymax", "grd_zmin", "grd_zmax",
                  "xspac"):
            arr = getattr(d, k)
            for b in range(self.ngrd):
                arr[b] = float(self.a[k][b])
        for k in _SCALARS:
            setattr(d, k, float(self.s[k]))
        d.isotropic = 1 if self.s["isotropic"] else 0
        d.tilting = 1 if self.s.get("tilting") else 0
        for k in ("tilt", "dtilt", "t_equiv", "planet_per", "xdip", "ydip", "zdip", "r_inner"):
            setattr(d, k, float(self.s.get(k, 0.0)))
        d.ut = float(self.s.get("ut", 0.0))
        d.dip_b0 = float(self.s["b0"])
        self._desc = d
        self.h = L.g3m3_run_create(_ct.byref(d))
        if not self.h:
            raise _lib.G3M3Error("g3m3_run_create failed")
        self._L = L

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self._L.g3m3_run_destroy(_ct.c_void_p(h))
            self.h = None

    def _pull(self):
        n = self.ngrd
        t = _ct.c_double(0)
        ts = (_ct.c_double * n)()
        tsn = (_ct.c_double * n)()
        to = (_ct.c_double * n)()
        tn = (_ct.c_double * n)()
        cu = _ct.c_longlong(0)
        bs = _ct.c_longlong(0)
        self._L.g3m3_run_info(_ct.c_void_p(self.h), _ct.byref(t), ts, tsn, to, tn,
                              _ct.byref(cu), _ct.byref(bs))
        self.t = t.value
        self.t_step[:] = list(ts)
        self.t_stepnew[:] = list(tsn)
        self.t_old[:] = list(to)
        self.t_new[:] = list(tn)
        self.cell_updates = cu.value
        self.box_steps = bs.value

    def init_dt(self):
        if self._L.g3m3_run_init_dt(_ct.c_void_p(self.h)):
            raise _lib.G3M3Error("g3m3_run_init_dt failed")
        self._pull()

    def full_step(self, dt_override=None):
        arg = None
        if dt_override is not None:
            arg = (_ct.c_double * self.ngrd)(*[float(x) for x in dt_override])
        if self._L.g3m3_run_full_step(_ct.c_void_p(self.h), arg):
            raise _lib.G3M3Error("g3m3_run_full_step failed")
        self._pull()

    def step_log(self):
        """The box-steps the last full step took, in order: (box, t_old,
        t_new) after each, as speeds.dat records them."""
        L = self._L
        L.g3m3_run_step_log.restype = _ct.c_int
        n = L.g3m3_run_step_log(_ct.c_void_p(self.h), 0, None, None, None)
        box = (_ct.c_int * max(n, 1))()
        to = (_ct.c_double * max(n, 1))()
        tn = (_ct.c_double * max(n, 1))()
        L.g3m3_run_step_log(_ct.c_void_p(self.h), n, box, to, tn)
        return [(int(box[q]), float(to[q]), float(tn[q])) for q in range(n)]

    def set_stream(self, stream):
        """Issue this run's kernels on `stream` (a torch.cuda.Stream)."""
        self._L.g3m3_run_set_stream(_ct.c_void_p(self.h), _ct.c_void_p(stream.cuda_stream))
        self._stream = stream

    def full_step_async(self, dt_override=None):
        """Issue a full step's kernels and return (finish() completes it)."""
        arg = None
        if dt_override is not None:
            arg = (_ct.c_double * self.ngrd)(*[float(x) for x in dt_override])
        if self._L.g3m3_run_full_step_async(_ct.c_void_p(self.h), arg):
            raise _lib.G3M3Error("g3m3_run_full_step_async failed")

    def finish(self):
        if self._L.g3m3_run_finish(_ct.c_void_p(self.h)):
            raise _lib.G3M3Error("g3m3_run_finish failed")
        self._pull()

    def graph_steps(self, nsteps: int):
        """Issue `nsteps` full steps of a single-grid run as replays of a
        captured HIP graph (time step kept on the device); graph_finish()
        waits for them."""
        if self._L.g3m3_run_graph_steps(_ct.c_void_p(self.h), int(nsteps)):
            raise _lib.G3M3Error("g3m3_run_graph_steps failed")

    def fct_shape(self):
        """The FCT launch shape chosen by the run's autotuner: (k-planes per
        block, block order 'xcd' or 'plain', {candidate: microseconds})."""
        kc, rm = _ct.c_int(0), _ct.c_int(0)
        us = (_ct.c_float * 6)()
        self._L.g3m3_run_fct_shape(_ct.c_void_p(self.h), _ct.byref(kc), _ct.byref(rm), us)
        cands = {f"{d}/{o}": round(float(us[2 * i + j]), 2)
                 for i, d in enumerate((8, 16, 32)) for j, o in enumerate(("xcd", "plain"))}
        return kc.value, ("plain" if rm.value else "xcd"), cands

    def set_fct_shape(self, kc_len: int, order: str = "xcd"):
        if self._L.g3m3_run_set_fct_shape(_ct.c_void_p(self.h), int(kc_len),
                                          int(order == "plain")):
            raise _lib.G3M3Error(f"bad FCT shape {kc_len}")

    def species_kernels(self):
        """The species push and Lapidus kernels chosen by the run's
        autotuner: ('march' or 'cell', {candidate: microseconds of one push
        plus one Lapidus pass})."""
        m = _ct.c_int(0)
        us = (_ct.c_float * 2)()
        self._L.g3m3_run_species_kernels(_ct.c_void_p(self.h), _ct.byref(m), us)
        return ("march" if m.value else "cell"), {"cell": round(float(us[0]), 2),
                                                   "march": round(float(us[1]), 2)}

    def set_species_kernels(self, kind: str):
        if kind not in ("march", "cell") or self._L.g3m3_run_set_species_kernels(
                _ct.c_void_p(self.h), int(kind == "march")):
            raise _lib.G3M3Error(f"bad species kernel family {kind!r}")

    def graph_finish(self):
        if self._L.g3m3_run_graph_finish(_ct.c_void_p(self.h)):
            raise _lib.G3M3Error("g3m3_run_graph_finish failed")
        self._pull()

    def sync_arrays(self):
        """Put every box's state back in the state tensors and its old state
        in the old tensors (the runtime swaps the two per box-step)."""
        if self._L.g3m3_run_sync_arrays(_ct.c_void_p(self.h)):
            raise _lib.G3M3Error("g3m3_run_sync_arrays failed")

    def divb(self, box, errlim, mode=0, max_iters=0):
        self.sync_arrays()
        return super().divb(box, errlim, mode, max_iters)

    # -- fluid (restart) files in the reference's format -------------------------
    @staticmethod
    def _host(a):
        return a.cpu().numpy() if hasattr(a, "cpu") else a

    def _fluid_arrays(self):
        return [self.a[k] for k in FLUID_ORDER]

    def _list_caps(self):
        return tuple(int(self.a["ij" + nm].shape[0]) for nm in ("srf", "mid", "zero"))

    def write_fluid(self, path):
        """Write the run (t, the 37 arrays, the planet lists) as a fluid file
        the reference reads at restart (code/3d_main.f90:972-1011)."""
        self.sync_arrays()
        host = [t.cpu().numpy() for t in self._fluid_arrays()]
        lists = {k: np.ascontiguousarray(self._host(self.a[k])) for k in _LIST_KEYS}
        msrf, mmid, mzero = self._list_caps()
        rc = _lib.call("fluid_write", path.encode() if isinstance(path, str) else path,
                       _lib.F64(self.t), self.nx, self.ny, self.nz, self.ngrd,
                       (_ct.c_void_p * len(host))(*[h.ctypes.data for h in host]),
                       int(self.s["mbndry"]), msrf, mmid, mzero,
                       lists["parm_srf"], lists["parm_mid"], lists["parm_zero"],
                       lists["ijsrf"], lists["ijmid"], lists["ijzero"],
                       lists["numsrf"], lists["nummid"], lists["numzero"])
        if rc:
            raise _lib.G3M3Error(f"writing {path} failed ({rc})")

    def read_fluid(self, path):
        """Restart from a fluid file: t, the 37 arrays and the planet lists
        replace the run's; old and wrk are zeroed, as in a fresh reference
        process, and the first time step is found again (init_dt)."""
        torch = _torch()
        self.sync_arrays()
        host = [np.empty(tuple(t.shape), np.float32) for t in self._fluid_arrays()]
        lists = {k: np.empty(tuple(self.a[k].shape), self._host(self.a[k]).dtype)
                 for k in _LIST_KEYS}
        msrf, mmid, mzero = self._list_caps()
        t = _ct.c_double(0.0)
        rc = _lib.call("fluid_read", path.encode() if isinstance(path, str) else path, t,
                       self.nx, self.ny, self.nz, self.ngrd,
                       (_ct.c_void_p * len(host))(*[h.ctypes.data for h in host]),
                       int(self.s["mbndry"]), msrf, mmid, mzero,
                       lists["parm_srf"], lists["parm_mid"], lists["parm_zero"],
                       lists["ijsrf"], lists["ijmid"], lists["ijzero"],
                       lists["numsrf"], lists["nummid"], lists["numzero"])
        if rc:
            raise _lib.G3M3Error(f"reading {path} failed ({rc})")
        for k in ("numsrf", "nummid", "numzero"):
            if not np.array_equal(lists[k], self._host(self.a[k])):
                raise _lib.G3M3Error(f"{path}: planet list {k} differs from the run's; "
                                     "create the run from the file's lists")
        for dev, h in zip(self._fluid_arrays(), host):
            dev.copy_(torch.from_numpy(h))
        for k in _LIST_KEYS:
            if hasattr(self.a[k], "copy_"):
                self.a[k].copy_(torch.from_numpy(lists[k]))
            else:
                self.a[k][...] = lists[k]
        for k in FIELDS:
            self.a["old" + k].zero_()
            self.a["wrk" + k].zero_()
        torch.cuda.synchronize()
        self._L.g3m3_run_set_time(_ct.c_void_p(self.h), _ct.c_double(t.value))
        self.init_dt()

    def state_host(self, prefix=""):
        self.sync_arrays()
        return super().state_host(prefix)


# -----------------------------------------------------------------------------
# Many independent runs advanced together: each run's full step is issued on
# one of a few HIP streams, then all are finished -- small boxes that leave
# most of the chip idle one at a time overlap on the GPU.


class Batch:
    """Independent runs (any sizes) stepped together on `streams` streams."""

    def __init__(self, workloads, streams: int = 8):
        torch = _torch()
        self.runs = [FastRun(wl) for wl in workloads]
        torch.cuda.synchronize()   # uploads done before other streams touch the arrays
        self.streams = [torch.cuda.Stream() for _ in range(max(1, streams))]
        for i, r in enumerate(self.runs):
            r.set_stream(self.streams[i % len(self.streams)])

    def init_dt(self):
        for r in self.runs:
            r.init_dt()

    def full_step(self):
        for r in self.runs:
            r.full_step_async()
        for r in self.runs:
            r.finish()

    def graph_steps(self, nsteps: int):
        """`nsteps` full steps of every run (single-grid runs), each run
        replaying its captured step graph on its stream with no host wait
        between steps; returns when all are done."""
        for r in self.runs:
            r.graph_steps(nsteps)
        for r in self.runs:
            r.graph_finish()

    def fused_steps(self, nsteps: int, groups: int = 4):
        """`nsteps` full steps of every run (single-grid runs) in lockstep
        groups: the runs are dealt into `groups` groups of about equal cells
        (largest first, in snake order), each group steps with one launch
        per stage for all its runs -- a grid concatenating the runs' own
        grids (g3m3_batch_*) -- replayed from one captured graph per step on
        the group's own stream, so one group's small stages overlap another's
        large ones; returns when all are done.  Each run's results equal its
        own graph_steps' bit for bit."""
        L = _lib.lib()
        if getattr(self, "_fused", None) is None:
            L.g3m3_batch_create.restype = _ct.c_void_p
            for nm in ("g3m3_batch_steps", "g3m3_batch_finish", "g3m3_batch_launches"):
                getattr(L, nm).restype = _ct.c_int
            torch = _torch()
            torch.cuda.synchronize()
            ng = max(1, min(int(groups), len(self.runs)))
            order = sorted(range(len(self.runs)), key=lambda i: -self.runs[i].nx * self.runs[i].ny
                           * self.runs[i].nz)
            members = [[] for _ in range(ng)]
            for rank, i in enumerate(order):
                g = rank % (2 * ng)
                members[g if g < ng else 2 * ng - 1 - g].append(i)
            self._fused, self._fused_streams = [], []
            for m in members:
                if not m:
                    continue
                st = torch.cuda.Stream()
                hs = (_ct.c_void_p * len(m))(*[self.runs[i].h for i in sorted(m)])
                h = L.g3m3_batch_create(hs, len(m), _ct.c_void_p(st.cuda_stream))
                if not h:
                    raise _lib.G3M3Error("g3m3_batch_create failed")
                self._fused.append(h)
                self._fused_streams.append(st)
        for h in self._fused:
            if L.g3m3_batch_steps(_ct.c_void_p(h), int(nsteps)):
                raise _lib.G3M3Error("g3m3_batch_steps failed")
        for h in self._fused:
            if L.g3m3_batch_finish(_ct.c_void_p(h)):
                raise _lib.G3M3Error("g3m3_batch_finish failed")
        for r in self.runs:
            r._pull()

    @property
    def fused_launches(self) -> int:
        """Multi-run launches per lockstep step, summed over the groups (0
        before fused_steps)."""
        hs = getattr(self, "_fused", None) or []
        return sum(int(_lib.lib().g3m3_batch_launches(_ct.c_void_p(h))) for h in hs)

    def __del__(self):
        hs = getattr(self, "_fused", None)
        L = getattr(_lib, "_lib", None)     # None once the interpreter tears the module down
        if hs and L is not None:
            for h in hs:
                L.g3m3_batch_destroy(_ct.c_void_p(h))
            self._fused = None

    @property
    def cell_updates(self) -> int:
        return sum(r.cell_updates for r in self.runs)
