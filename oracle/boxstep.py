"""The reference's per-box-step and full-step call sequence, on the CPU.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

``OracleRun`` replays code/3d_main.f90:1925-3095 -- dt synchronisation,
subcycling, the 2-step Lax-Wendroff push, Lapidus, FCT, boundaries,
nested-grid sync -- calling either the NumPy restatement
(``backend="restate"``) or the reference's own compiled subroutines
(``backend="ref"``/``"ref_omp"``) with the reference's argument lists.
Quirks kept (they change results): the step-1 ``push_elec`` writes the
state ``epres`` and the electron fluxes use the previous box-step's
``wrkepres`` (:2271, :2517); inner boxes apply their step-2 and Lapidus
flanks to the ``wrk`` arrays and reuse ``t_grid`` (:2589-2621, :2797);
``wrk`` and ``old`` start zeroed (``-zero -save``).

Extension (no reference counterpart): ``eta`` adds ``eta*J/n`` to the
interior electric field of every box after BandE's list term, then the
flank copies are redone; with ``eta = 0`` the step is the reference's.
"""
from __future__ import annotations

import numpy as np

from . import argnames as A
from . import restate as O

F = np.float32
SPECIES = "qho"
SP_FIELDS = ("rho", "presx", "presy", "presz", "px", "py", "pz",
             "presxy", "presxz", "presyz")
FIELDS = [s + f for s in SPECIES for f in SP_FIELDS] + ["epres", "bx", "by", "bz"]


class OracleRun:
    def __init__(self, wl, backend: str = "restate", threads: int = 1):
        self.wl = wl
        self.nx, self.ny, self.nz, self.ngrd = wl.nx, wl.ny, wl.nz, wl.ngrd
        self.backend = backend
        if backend.startswith("ref"):
            from . import ref as R
            self.lib = R.RefLib(omp=(backend == "ref_omp"))
            if backend == "ref_omp":
                self.lib.set_threads(threads)
            rot = self.lib.rotation()
            rot.v_rot = float(wl.scal["v_rot"])
            rot.r_rot = float(wl.scal["r_rot"])
        a = {}
        for k in FIELDS:
            a[k] = wl.fields[k].copy()
            a["old" + k] = np.zeros_like(a[k])
            a["wrk" + k] = np.zeros_like(a[k])
        for k in ("bx0", "by0", "bz0", "rst"):
            a[k] = wl.fields[k].copy()
        shp3 = (self.nz, self.ny, self.nx)
        for k in ("curx", "cury", "curz", "evx", "evy", "evz", "tvx", "tvy", "tvz",
                  "efldx", "efldy", "efldz", "bsx", "bsy", "bsz", "btot",
                  "vvx", "vvy", "vvz"):
            a[k] = np.zeros(shp3, F)
        for k in ("pxmax", "pymax", "pzmax", "csmax", "alfmax", "fastest"):
            a[k] = np.zeros(1, F)
        for k, v in wl.grid.items():
            a[k] = v.copy()
        for k, v in wl.lists.items():
            if isinstance(v, np.ndarray):
                a[k] = v
        n = self.ngrd
        a["t_old"] = np.zeros(n, np.float64)
        a["t_new"] = np.zeros(n, np.float64)
        self.a = a
        s = dict(wl.scal)
        s.update(nx=self.nx, ny=self.ny, nz=self.nz, n_grids=n,
                 msrf=wl.lists["msrf"], mmid=wl.lists["mmid"], mzero=wl.lists["mzero"])
        self.s = s
        self.t_step = np.zeros(n, np.float64)
        self.t_stepnew = np.zeros(n, np.float64)
        self.t = 0.0
        # the main loop's tilt and universal time (code/3d_main.f90:1966-1977)
        self.tilt = self.old_tilt = F(s.get("tilt", 0.0))
        self.ut = self.utold = float(s.get("ut", 0.0))
        self.cell_updates = 0
        self.box_steps = 0
        self.dt_log = []
        self.step_log = []    # (box, t_old, t_new) per box-step, as speeds.dat

    # -- plumbing ------------------------------------------------------------
    def args(self, names):
        out = []
        for nm in names:
            if isinstance(nm, str):
                out.append(self.a[nm] if nm in self.a else self.s[nm])
            else:
                out.append(nm)
        return out

    def call(self, name, names, **kw):
        args = self.args(names)
        if self.backend == "restate":
            getattr(O, name)(*args, **kw)
        else:
            self.lib.call(name, *args)

    # -- pieces of the box-step ------------------------------------------------
    def _tilt(self, box, dt_off):
        """The box's b0 for the tilt and rotation at t_old + dt_off: the half
        step before step 1 (code/3d_main.f90:2173-2187), the full step
        before step 2 (:2465-2478)."""
        s, a = self.s, self.a
        sin_tilt, cos_tilt, rot_angle = O.tilt_rotation(
            self.old_tilt, self.utold, a["t_old"][box - 1], dt_off, s["dtilt"], s["t_equiv"],
            s["planet_per"])
        rot = dict(rot_angle=rot_angle, xdip=s["xdip"], ydip=s["ydip"], zdip=s["zdip"],
                   sin_tilt=sin_tilt, cos_tilt=cos_tilt, b0=s["b0"])
        args = [a["bx0"], a["by0"], a["bz0"], self.nx, self.ny, self.nz, self.ngrd,
                s["mbndry"], box, a["ijzero"], a["numzero"], s["mzero"], F(s["r_inner"]),
                a["grd_xmin"], a["grd_xmax"], a["grd_ymin"], a["grd_ymax"], a["grd_zmin"],
                a["grd_zmax"]]
        if self.backend == "restate":
            O.mak_dip_grd(*args, **rot)
        else:
            r = self.lib.rotation()
            for k, v in rot.items():
                setattr(r, k, float(v))
            self.lib.call("mak_dip_grd", *args)

    def _btot(self, pre, b):
        a = self.a
        a["bsx"][...] = a["bx0"][b] + a[pre + "bx"][b]
        a["bsy"][...] = a["by0"][b] + a[pre + "by"][b]
        a["bsz"][...] = a["bz0"][b] + a[pre + "bz"][b]
        a["btot"][...] = np.maximum(np.sqrt(a["bsx"] * a["bsx"] + a["bsy"] * a["bsy"]
                                            + a["bsz"] * a["bsz"]), self.s["smallbit"])

    def _efield(self, pre, box):
        s = self.s
        self.call("calcur", [pre + "bx", pre + "by", pre + "bz", "nx", "ny", "nz", "n_grids",
                             "box", "curx", "cury", "curz", "rx", "ry", "rz"])
        self._btot(pre, box - 1)
        self.call("fnd_evel", [pre + "qpx", pre + "qpy", pre + "qpz", pre + "qrho",
                               pre + "hpx", pre + "hpy", pre + "hpz", pre + "hrho",
                               pre + "opx", pre + "opy", pre + "opz", pre + "orho",
                               "curx", "cury", "curz", "evx", "evy", "evz", "tvx", "tvy", "tvz",
                               "nx", "ny", "nz", "n_grids", "box", "rmassq", "rmassh", "rmasso",
                               "reynolds"])
        self.call("BandE", ["efldx", "efldy", "efldz", "bsx", "bsy", "bsz", "curx", "cury",
                            "curz", "evx", "evy", "evz", "btot", pre + "epres", pre + "qrho",
                            pre + "hrho", pre + "orho", "rst", "resist", "reynolds",
                            "nx", "ny", "nz", "n_grids", "box", "rmassq", "rmassh", "rmasso",
                            "ijmid", "nummid", "ijzero", "mbndry", "numzero", "mmid", "mzero",
                            "rx", "ry", "rz"])
        eta = F(s["eta"])
        if eta != 0:
            a, b, C = self.a, box - 1, O.C
            q, h, o = a[pre + "qrho"][b][C], a[pre + "hrho"][b][C], a[pre + "orho"][b][C]
            eden = q / s["rmassq"] + h / s["rmassh"] + o / s["rmasso"]
            for e, c in (("efldx", "curx"), ("efldy", "cury"), ("efldz", "curz")):
                a[e][C] = a[e][C] + eta * a[c][C] / eden
            for e in ("efldx", "efldy", "efldz"):
                E = a[e]
                E[:, :, 0] = E[:, :, 1]
                E[:, :, -1] = E[:, :, -2]
            for e in ("efldx", "efldy", "efldz"):
                E = a[e]
                E[0] = E[1]
                E[-1] = E[-2]
            for e in ("efldx", "efldy", "efldz"):
                E = a[e]
                E[:, 0, :] = E[:, 1, :]
                E[:, -1, :] = E[:, -2, :]

    def _push(self, out, old, src, delt):
        self.s["delt"] = F(delt)
        self.call("push_elec", ["wrkepres", "oldepres", "epres", "evx", "evy", "evz", "gamma",
                                "gamma1", "nx", "ny", "nz", "n_grids", "box", "delt",
                                "rx", "ry", "rz"])
        for sp, rm in zip(SPECIES, ("rmassq", "rmassh", "rmasso")):
            self.call("push_ion", A.push_ion_names(out, old, src, sp, rm))
        self.call("push_bfld", [out + "bx", out + "by", out + "bz", old + "bx", old + "by",
                                old + "bz", "efldx", "efldy", "efldz", "nx", "ny", "nz",
                                "n_grids", "box", "delt", "rx", "ry", "rz"])

    def _bounds(self, pre, box, flank_target, flank_source):
        if box == self.ngrd:
            self.call("bndry_outer", A.bndry_outer_names(pre))
        else:
            names = A.flanks_names(flank_target, flank_source, "old")
            names[names.index("work")] = "vvx"
            self.call("bndry_flanks", names)
        if box <= self.s["mbndry"]:
            self.call("bndry_inner", A.bndry_inner_names(pre),
                      **({"v_rot": self.s["v_rot"], "r_rot": self.s["r_rot"]}
                         if self.backend == "restate" else {}))

    def _validate(self, pre, box):
        self.call("set_rho", A.set_rho_names(pre))
        self.s["vlim"] = F(self.s["vlim_factor"]) * F(box)
        self.call("set_speed_agrd", A.set_speed_names(pre))

    def _set_t(self, t):
        from .ref import F64
        self.s["t"] = F64(t) if self.backend.startswith("ref") else t

    # -- code/3d_main.f90:2163-3065 ---------------------------------------------
    def box_step(self, box):
        a, s = self.a, self.s
        b = box - 1
        a["t_old"][b] = a["t_new"][b]
        a["t_new"][b] = a["t_old"][b] + self.t_step[b]
        self.step_log.append((box, float(a["t_old"][b]), float(a["t_new"][b])))
        delt = F(self.t_step[b])
        delt2 = delt / F(2.)
        s["box"] = box
        if s.get("tilting"):                          # :2173-2187
            self._tilt(box, delt2)
        for k in FIELDS:                              # store_array x34
            a["old" + k][b] = a[k][b]
        s["rx"] = s["ry"] = s["rz"] = F(a["xspac"][b])
        # step 1 (:2243-2458)
        self._efield("", box)
        self._push("wrk", "old", "", F(0.5) * delt)
        if box < self.ngrd:
            tg = a["t_old"][b] + float(delt2)
            tg = a["t_new"][b + 1] if tg > a["t_new"][b + 1] else tg
            tg = a["t_old"][b + 1] if tg < a["t_old"][b + 1] else tg
            self._set_t(tg)
        self._bounds("wrk", box, "wrk", "")
        self._validate("wrk", box)
        # step 2 (:2465-2684)
        if s.get("tilting"):                          # :2465-2478
            self._tilt(box, delt)
        self._efield("wrk", box)
        self._push("", "old", "wrk", delt)
        if box < self.ngrd:
            tg = a["t_old"][b] + float(delt)
            tg = a["t_new"][b + 1] if tg > a["t_new"][b + 1] else tg
            tg = a["t_old"][b + 1] if tg < a["t_old"][b + 1] else tg
            self._set_t(tg)
        self._bounds("", box, "wrk", "")
        self._validate("", box)
        # Lapidus (:2692-2890)
        self.s["delt"] = delt
        sb = s["smallbit"]
        for sp in SPECIES:
            den = np.maximum(a[sp + "rho"][b], sb)
            a["curx"][...] = a[sp + "px"][b] / den
            a["cury"][...] = a[sp + "py"][b] / den
            a["curz"][...] = a[sp + "pz"][b] / den
            names = A.lap_plasma_names(sp, "", "wrk")
            names[names.index("vvx"):names.index("vvx") + 3] = ["curx", "cury", "curz"]
            self.call("lap_plasma", names)
        self.call("fnd_vtot", ["qpx", "qpy", "qpz", "qrho", "hpx", "hpy", "hpz", "hrho",
                               "opx", "opy", "opz", "orho", "curx", "cury", "curz",
                               "nx", "ny", "nz", "n_grids", "box", "rmassq", "rmassh", "rmasso"])
        self.call("lap_elec", ["epres", "wrkepres", "curx", "cury", "curz", "nx", "ny", "nz",
                               "n_grids", "box", "chirho", "chipxyz", "chierg", "delt"])
        self.call("lap_bfld", ["bx", "by", "bz", "wrkbx", "wrkby", "wrkbz", "curx", "cury",
                               "curz", "nx", "ny", "nz", "n_grids", "box", "chirho",
                               "chipxyz", "chierg", "delt"])
        self._bounds("wrk", box, "wrk", "")
        self._validate("wrk", box)
        # FCT (:2896-3030)
        self.call("flux_correct", A.flux_correct_names())
        self._bounds("", box, "", "")
        self._validate("", box)
        fastest = a["fastest"][0]
        self.t_stepnew[b] = min(F(s["stepsz"]) * F(a["xspac"][b]) / fastest,
                                F(s["t_step_max"]))
        if box < self.ngrd:
            if abs(a["t_new"][b] - a["t_new"][b + 1]) <= 0.005 * self.t_step[b]:
                self.call("bndry_corer", A.corer_names(""))
                a["t_old"][b + 1] = a["t_new"][b + 1]
        self.cell_updates += self.nx * self.ny * self.nz
        self.box_steps += 1

    # -- code/3d_main.f90:1861-1903 ---------------------------------------------
    def init_dt(self):
        a, s = self.a, self.s
        for box in range(self.ngrd, 0, -1):
            b = box - 1
            a["t_old"][b] = 0.
            a["t_new"][b] = 0.
            s["box"] = box
            self._validate("", box)
            self.t_stepnew[b] = min(F(s["stepsz"]) * F(a["xspac"][b]) / a["fastest"][0],
                                    F(s["t_step_max"]))

    # -- code/3d_main.f90:1925-3095 ---------------------------------------------
    def full_step(self, dt_override=None):
        a, n = self.a, self.ngrd
        ts = self.t_step
        smallest = F(0)
        m_small = n
        for box in range(n, 0, -1):
            b = box - 1
            ts[b] = self.t_stepnew[b]
            a["t_old"][b] = 0.
            a["t_new"][b] = 0.
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
            raise FloatingPointError(f"time step {float(dm)} below t_step_min 1e-6 or not finite")
        m_step = int(ts[n - 1] / ts[m_small - 1])
        delt = F(ts[n - 1])
        self.dt_log.append(ts.copy())
        self.t, self.ut, self.tilt, self.utold, self.old_tilt = O.main_clock(
            self.t, self.ut, self.tilt, delt, self.s.get("t_equiv", 1.0),
            self.s.get("dtilt", 0.0))
        for _ in range(m_step):
            for box in range(n, 0, -1):
                b = box - 1
                yes = box == 1 or (a["t_old"][b] == a["t_new"][b]
                                   and a["t_new"][b] < ts[n - 1]
                                   and abs(a["t_new"][b] - a["t_new"][0]) <= 0.005 * ts[0])
                if yes:
                    self.box_step(box)
        a["t_old"][0] = a["t_new"][0]
        for box in range(1, n):
            self.s["box"] = box
            self.call("bndry_corer", A.corer_names(""))
            self.call("set_rho", A.set_rho_names(""))

    def state(self):
        return {k: self.a[k] for k in FIELDS}
