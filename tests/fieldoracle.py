"""Test infrastructure: the frozen-plasma field step (g3m3_field_model_step)
composed from the oracle's restatements of the reference subroutines, in
float32, in the order of code/3d_main.f90 for one outermost box
(box = n_grids = 1, mbndry = 0):

  stage 1 (:2243-2331): calcur, bs = b0 + b, fnd_evel, BandE (+ eta J/n on
      the interior, added before BandE's face copies: the copies are redone
      after it, which gives the same as adding it first), push_bfld with
      0.5*delt from old = b into a copy of b, bndry_outer;
  stage 2 (:2486-2576): the same from the stage-1 field, push_bfld with delt;
  lap_bfld (:2765) into a copy of the stage-1 field, bndry_outer (:2774);
  fcsmooth with diferg (code/sub_flux_correct.f90:183), bndry_outer (:2932).
"""
from __future__ import annotations

import numpy as np

from oracle import restate as R

F = np.float32


class FieldOracle:
    def __init__(self, setup):
        s = setup
        self.s = s
        self.nx, self.ny, self.nz = s.nx, s.ny, s.nz
        shp = (1, s.nz, s.ny, s.nx)
        self.pl = {k: v.reshape(shp).astype(F) for k, v in s.plasma.items()}
        self.b0 = [s.b0[c].reshape(shp).astype(F) for c in range(3)]
        self.sc = s.scal
        z = lambda: np.zeros((s.nz, s.ny, s.nx), F)
        self.tv = [z(), z(), z()]
        self.vt = [z(), z(), z()]
        p = self.pl
        R.fnd_vtot(p["qpx"], p["qpy"], p["qpz"], p["qrho"], p["hpx"], p["hpy"], p["hpz"],
                   p["hrho"], p["opx"], p["opy"], p["opz"], p["orho"], *self.vt,
                   s.nx, s.ny, s.nz, 1, 1, self.sc["rmassq"], self.sc["rmassh"],
                   self.sc["rmasso"])
        # dummy plasma arrays for bndry_outer (only the field rules matter)
        self._dummy = [np.zeros(shp, F) for _ in range(31)]

    def _efield(self, b, eta):
        s, sc, p = self.s, self.sc, self.pl
        nx, ny, nz = s.nx, s.ny, s.nz
        J = [np.zeros((nz, ny, nx), F) for _ in range(3)]
        R.calcur(b[0], b[1], b[2], nx, ny, nz, 1, 1, *J, sc["rx"], sc["rx"], sc["rx"])
        bs = [(self.b0[c][0] + b[c][0]).astype(F) for c in range(3)]
        ev = [np.zeros((nz, ny, nx), F) for _ in range(3)]
        R.fnd_evel(p["qpx"], p["qpy"], p["qpz"], p["qrho"], p["hpx"], p["hpy"], p["hpz"],
                   p["hrho"], p["opx"], p["opy"], p["opz"], p["orho"], *J, *ev, *self.tv,
                   nx, ny, nz, 1, 1, sc["rmassq"], sc["rmassh"], sc["rmasso"], sc["reynolds"])
        E = [np.zeros((nz, ny, nx), F) for _ in range(3)]
        btot = np.zeros((nz, ny, nx), F)
        none = np.zeros((1, 3, 1), np.int32)
        cnt = np.zeros(1, np.int32)
        R.BandE(*E, *bs, *J, *ev, btot, p["epres"], p["qrho"], p["hrho"], p["orho"],
                np.zeros((1, nz, ny, nx), F), F(0), sc["reynolds"], nx, ny, nz, 1, 1,
                sc["rmassq"], sc["rmassh"], sc["rmasso"], none, cnt, none, 0, cnt, 1, 1,
                sc["rx"], sc["rx"], sc["rx"])
        # eta J / n on the interior, then BandE's face copies again
        q, h, o = p["qrho"][0], p["hrho"][0], p["orho"][0]
        eden = q / sc["rmassq"] + h / sc["rmassh"] + o / sc["rmasso"]
        C = (slice(1, -1),) * 3
        for e, j in zip(E, J):
            e[C] = e[C] + F(eta) * j[C] / eden[C]
        for e in E:
            e[:, :, 0] = e[:, :, 1]
            e[:, :, -1] = e[:, :, -2]
        for e in E:
            e[0] = e[1]
            e[-1] = e[-2]
        for e in E:
            e[:, 0, :] = e[:, 1, :]
            e[:, -1, :] = e[:, -2, :]
        return E

    def _outer(self, b, ctl):
        s = self.s
        d = self._dummy
        R.bndry_outer(*d[0:21], d[21], *d[22:31], self.sc["rmassq"], self.sc["rmassh"],
                      self.sc["rmasso"], b[0], b[1], b[2], *self.b0, s.nx, s.ny, s.nz, 1,
                      F(1), F(0), F(0), F(1), F(0), F(0), F(0), F(ctl[1]), F(ctl[2]),
                      F(ctl[3]), F(1), False)

    def step(self, b_np, ctl):
        """One step from b (3, nz, ny, nx) float32; returns the new b."""
        s, sc = self.s, self.sc
        nx, ny, nz = s.nx, s.ny, s.nz
        shp = (1, nz, ny, nx)
        b = [np.array(b_np[c], F).reshape(shp) for c in range(3)]
        eta = F(ctl[0])
        E = self._efield(b, eta)
        w1 = [x.copy() for x in b]
        R.push_bfld(*w1, *b, *E, nx, ny, nz, 1, 1, F(0.5) * F(sc["delt"]), sc["rx"], sc["rx"],
                    sc["rx"])
        self._outer(w1, ctl)
        E = self._efield(w1, eta)
        b2 = [x.copy() for x in b]
        R.push_bfld(*b2, *b, *E, nx, ny, nz, 1, 1, sc["delt"], sc["rx"], sc["rx"], sc["rx"])
        self._outer(b2, ctl)
        W = [x.copy() for x in w1]
        R.lap_bfld(*b2, *W, *self.vt, nx, ny, nz, 1, 1, F(0), F(0), sc["chierg"], sc["delt"])
        self._outer(W, ctl)
        out = [np.zeros(shp, F) for _ in range(3)]
        t = [np.zeros((nz, ny, nx), F) for _ in range(3)]
        for c in range(3):
            R.fcsmooth(out[c], b[c], W[c], nx, ny, nz, 1, 1, sc["diferg"], *t)
        self._outer(out, ctl)
        return np.stack([o[0] for o in out])

    def sample(self, b_np, pos):
        g = self.s.grid
        fields = [np.asarray(b_np[c], F).reshape(1, self.nz, self.ny, self.nx) for c in range(3)]
        out, gp = R.sample_fields(np.asarray(pos, F), 1, g["grid_minvals"], g["grid_maxvals"],
                                  g["xspac"], g["re_equiv"], fields)
        return out, gp
