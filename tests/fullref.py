"""Test infrastructure: a torch restatement of the whole box-step of an
outermost box (g3m3_full_model_step), written for torch.autograd.

Float32 runs reproduce the oracle's box-step (oracle/boxstep.py
OracleRun.box_step on a one-box run with mbndry = 0 and a fixed time step)
bit for bit, which pins the restatement; float64 runs give the
vector-Jacobian products and central differences the device adjoint
(g3m3_full_model_adjoint) is held to.  fcsmooth's limiter branches of a
float32 step can be recorded and taken by a float64 step (FieldRef.fcsmooth).

State: a list of 35 (nz, ny, nx) tensors, the 34 fields in the order
q, h, o x (rho, presx, presy, presz, px, py, pz, presxy, presxz, presyz),
epres, bx, by, bz, then the electron pressure of the last Lapidus pass.
ctl: 9 tensors (eta, sbx, sby, sbz, srho, spx, spy, spz, spress).

Reference call sequence: code/3d_main.f90:2243-2340 (stage 1), :2486-2580
(stage 2), :2698-2778 (Lapidus), :2896-2935 (flux correction), with
calcur code/sub_calcur.f90, fnd_evel code/sub_fnd_evel.f90, BandE
code/sub_BandE.f90, push_ion code/sub_push_ion.f90, push_elec
code/sub_push_elec.f90, push_bfld code/sub_push_bfld.f90, bndry_outer
code/sub_bndry_outer.f90:50-728, lap_plasma / lap_elec / lap_bfld
code/sub_lap_*.f90, fnd_vtot code/sub_fnd_vtot.f90, fcsmooth
code/sub_fcsmooth.f90.
"""
from __future__ import annotations

import numpy as np
import torch

from fieldref import FieldRef, _set
from ionref import _sqrt, lap_plasma, push_ion, species_velocity

from torch.overrides import TorchFunctionMode

EP, BX, WE = 30, 31, 34
# bndry_outer's rule per field of a species (code/sub_bndry_outer.f90)
RHO, PRES, PX, PY, PZ, OFF, EPK, BXK, BYK, BZK = range(10)
_SPECIES_KINDS = (RHO, PRES, PRES, PRES, PX, PY, PZ, OFF, OFF, OFF)
KINDS = list(_SPECIES_KINDS) * 3 + [EPK, BXK, BYK, BZK]
C3 = (slice(1, -1),) * 3


class FullRef(FieldRef):
    def __init__(self, setup, dtype=torch.float64):
        self.dt = dtype
        S = lambda x: torch.tensor(float(np.float32(x)), dtype=dtype)
        self.S = S
        self.s = setup
        self.b0 = tuple(torch.from_numpy(np.asarray(setup.b0[c], np.float32)).to(dtype)
                        for c in range(3))
        sc = setup.scal
        self.sc = {k: np.float32(v) for k, v in sc.items()}
        self.rx, self.delt = S(sc["rx"]), S(sc["delt"])
        self.rmq, self.rmh, self.rmo = S(sc["rmassq"]), S(sc["rmassh"]), S(sc["rmasso"])
        self.re, self.chierg, self.diferg = S(sc["reynolds"]), S(sc["chierg"]), S(sc["diferg"])
        self.difrho = S(sc["difrho"])
        self.gamma, self.gamma1 = S(sc["gamma"]), S(sc["gamma1"])
        self.smallbit = S(sc["smallbit"])
        self.two, self.three, self.six, self.half = S(2.0), S(3.0), S(6.0), S(0.5)

    def state(self, st_np):
        return [torch.from_numpy(np.asarray(st_np[i], np.float32)).to(self.dt)
                for i in range(len(st_np))]

    # --- inflow values of bndry_outer (code/sub_bndry_outer.f90:50-103) ---
    def inflow(self, ctl):
        S, sc = self.S, self.sc
        srho, spx, spy, spz, spress = ctl[4], ctl[5], ctl[6], ctl[7], ctl[8]
        rf = S(sc["rho_frac"])
        ofrac = S(np.float32(sc["rho_frac"] * sc["o_conc"]))
        frac_h = S(np.float32(sc["rmassh"] / sc["rmassq"]))
        frac_o = S(np.float32(sc["rmasso"] / sc["rmassq"]))
        half = self.half
        pq = half * spress
        z = torch.zeros((), dtype=self.dt)
        q = [srho, pq, pq, pq, spx, spy, spz, z, z, z]
        hp = half * spress * rf
        h = [srho * rf * frac_h, hp, hp, hp, spx * rf * frac_h, spy * rf * frac_h,
             spz * rf * frac_h, z, z, z]
        op = half * spress * ofrac
        o = [srho * ofrac * frac_o, op, op, op, spx * ofrac * frac_o, spy * ofrac * frac_o,
             spz * ofrac * frac_o, z, z, z]
        return q + h + o + [pq / S(sc["ti_te"]), ctl[1], ctl[2], ctl[3]]

    def outer_all(self, X, vals):
        """bndry_outer's eight passes, in order, on the 34 fields."""
        out = []
        for f in range(34):
            kind = KINDS[f]
            a = X[f].clone()
            if kind >= BXK:
                a[:, :, 0] = vals[f] - self.b0[kind - BXK][:, :, 0]
            else:
                a[:, :, 0] = vals[f]
            # (views are cloned before abs, which keeps its input for backward)
            v = a[:, :, -2].clone()
            a[:, :, -1] = torch.abs(v) if kind == PX else v
            lo, hi = a[:, 1, 1:].clone(), a[:, -2, 1:].clone()
            a[:, 0, 1:] = -torch.abs(lo) if kind == PY else lo
            a[:, -1, 1:] = torch.abs(hi) if kind == PY else hi
            v = a[-2, :, 1:].clone()
            a[-1, :, 1:] = torch.abs(v) if kind == PZ else v
            v = a[1, :, 1:].clone()
            a[0, :, 1:] = -torch.abs(v) if kind == PZ else v
            if kind == BXK:
                a[1:-1, 0, -1] = a[1:-1, 0, -2]
                a[1:-1, -1, -1] = a[1:-1, -1, -2]
            else:
                a[1:-1, 0, -1] = (a[1:-1, 0, -2] + a[1:-1, 1, -1]) / self.two
                a[1:-1, -1, -1] = (a[1:-1, -1, -2] + a[1:-1, -2, -1]) / self.two
            a[0, 1:-1, -1] = a[1, 1:-1, -1]
            if kind != EPK:
                a[-1, 1:-1, -1] = a[-1, 1:-1, -2]
            a[0, 0, 1:] = a[1, 0, 1:]
            a[0, -1, 1:] = a[1, -1, 1:]
            a[-1, 0, 1:] = a[-2, 0, 1:]
            a[-1, -1, 1:] = a[-2, -1, 1:]
            out.append(a)
        return out

    # --- the E-field of a set: calcur, total B, fnd_evel, BandE + eta J/n ---
    def efield_set(self, X, eta):
        I = self._i
        b = (X[BX], X[BX + 1], X[BX + 2])
        J = self.calcur(b)
        bs = tuple(self.b0[c] + b[c] for c in range(3))
        btot = torch.clamp(_sqrt(bs[0] * bs[0] + bs[1] * bs[1] + bs[2] * bs[2]),
                           min=float(self.sc["smallbit"]))
        q, h, o = X[0], X[10], X[20]
        d = self.S(0.000001)
        tot = (q + d) / self.rmq + (h + d) / self.rmh + (o + d) / self.rmo
        tden = q + h + o
        tv = tuple((X[4 + c] + X[14 + c] + X[24 + c]) / tden for c in range(3))
        ev = tuple(tv[c] - J[c] / tot / self.re for c in range(3))
        ex = -(ev[1] * bs[2] - ev[2] * bs[1])
        ey = -(ev[2] * bs[0] - ev[0] * bs[2])
        ez = -(ev[0] * bs[1] - ev[1] * bs[0])
        dxt = self.two * self.rx
        ep = X[EP]
        ar = I((q / self.rmq + h / self.rmh + o / self.rmo) * self.re)
        ex = _set(ex, C3, I(ex) - ((I(ep, 0, 0, 1) - I(ep, 0, 0, -1)) / dxt) / ar)
        ey = _set(ey, C3, I(ey) - ((I(ep, 0, 1) - I(ep, 0, -1)) / dxt) / ar)
        ez = _set(ez, C3, I(ez) - ((I(ep, 1) - I(ep, -1)) / dxt) / ar)
        en = I(q / self.rmq + h / self.rmh + o / self.rmo)
        E = [_set(e, C3, I(e) + eta * I(j) / en) for e, j in zip((ex, ey, ez), J)]
        out = []
        for e in E:
            e = e.clone()
            e[:, :, 0] = e[:, :, 1]
            e[:, :, -1] = e[:, :, -2]
            e[0] = e[1]
            e[-1] = e[-2]
            e[:, 0, :] = e[:, 1, :]
            e[:, -1, :] = e[:, -2, :]
            out.append(e)
        return dict(E=tuple(out), bs=bs, btot=btot, ev=ev, tv=tv)

    def push_elec(self, w, old, ev, delt):
        I = self._i
        dxt = self.two * self.rx
        evx, evy, evz = ev
        egx = (I(w, 0, 0, 1) - I(w, 0, 0, -1)) / dxt
        egy = (I(w, 0, 1) - I(w, 0, -1)) / dxt
        egz = (I(w, 1) - I(w, -1)) / dxt
        e = I(old) - delt * self.gamma * (
            ((I(w, 0, 0, 1) * I(evx, 0, 0, 1) - I(w, 0, 0, -1) * I(evx, 0, 0, -1)) / dxt)
            + ((I(w, 0, 1) * I(evy, 0, 1) - I(w, 0, -1) * I(evy, 0, -1)) / dxt)
            + ((I(w, 1) * I(evz, 1) - I(w, -1) * I(evz, -1)) / dxt))
        return _set(old, C3, e + delt * self.gamma1 * (I(evx) * egx + I(evy) * egy
                                                         + I(evz) * egz))

    def _push_stage(self, old, src, ef, delt32):
        """push_ion x3 and push_bfld into copies of old (interiors replaced)."""
        sc = self.sc
        grd = [sc["xmin"], sc["xmax"], sc["ymin"], sc["ymax"], sc["zmin"], sc["zmax"]]
        out = list(old[:34])
        for si, rm in enumerate(("rmassq", "rmassh", "rmasso")):
            blk = slice(10 * si, 10 * si + 10)
            p = dict(rmass=sc[rm], delt=delt32, grav=sc["grav"], re_equiv=sc["re_equiv"],
                     reynolds=sc["reynolds"], grd=grd)
            out[blk] = push_ion(old[blk], old[blk], src[blk], ef["bs"], ef["btot"], ef["E"],
                                ef["tv"], p, self.dt)
        out[BX:BX + 3] = self.push(tuple(old[BX:BX + 3]), ef["E"], self.S(delt32))
        return out

    def _lap1(self, q, w, v, chi):
        I = self._i
        vx, vy, vz = v
        uxp = torch.abs(I(vx, 0, 0, 1) - I(vx))
        uxm = torch.abs(I(vx) - I(vx, 0, 0, -1))
        uyp = torch.abs(I(vy, 0, 1) - I(vy))
        uym = torch.abs(I(vy) - I(vy, 0, -1))
        uzp = torch.abs(I(vz, 1) - I(vz))
        uzm = torch.abs(I(vz) - I(vz, -1))
        c = I(q)
        d = self.delt
        val = c + chi * (d * (uxp * (I(q, 0, 0, 1) - c) - uxm * (c - I(q, 0, 0, -1)))
                         + d * (uyp * (I(q, 0, 1) - c) - uym * (c - I(q, 0, -1)))
                         + d * (uzp * (I(q, 1) - c) - uzm * (c - I(q, -1))))
        return _set(w, C3, val)

    def step(self, st, ctl, codes=None, record=None, flank=None):
        """One box-step.  codes / record: fcsmooth's limiter branches per
        field (34 lists of three per-axis tensors), see FieldRef.fcsmooth.
        flank: an inner box's step -- dict(old=, cur=) the parent's two
        states (lists of >= 34 tensors), t12 the four time weights, geom
        from flank_geom -- whose stages are bounded by the flanks (step 1's
        set at (t1, t2) = t12[0:2], the Lapidus set and the output at
        t12[2:4]) and whose stage-2 set keeps the state's shell."""
        sc = self.sc
        S = list(st[:34])
        eta = ctl[0]
        vals = self.inflow(ctl)
        delt = sc["delt"]
        if flank is None:
            b1 = b3 = lambda X: self.outer_all(X, vals)
        else:
            t = [self.S(x) for x in flank["t12"]]
            b1 = lambda X: self.flanks(X, flank, t[0], t[1])
            b3 = lambda X: self.flanks(X, flank, t[2], t[3])
        # stage 1: push from S with delt/2; the pressure keeps the stale WE
        ef = self.efield_set(S, eta)
        W1 = self._push_stage(S, S, ef, np.float32(0.5) * delt)
        W1[EP] = st[WE]
        W1 = b1(W1)
        # stage 2
        ef = self.efield_set(W1, eta)
        S2 = self._push_stage(S, W1, ef, delt)
        S2[EP] = self.push_elec(W1[EP], S[EP], ef["ev"], self.delt)
        if flank is None:
            S2 = self.outer_all(S2, vals)
        # Lapidus into W1
        W3 = list(W1)
        for si in range(3):
            blk = slice(10 * si, 10 * si + 10)
            v = species_velocity(S2[10 * si], S2[10 * si + 4:10 * si + 7], sc["smallbit"],
                                 self.dt)
            W3[blk] = lap_plasma(S2[blk], W1[blk], v, sc["chirho"], sc["chipxyz"], delt,
                                 self.dt)
        tden = S2[0] + S2[10] + S2[20]
        vt = tuple((S2[4 + c] + S2[14 + c] + S2[24 + c]) / tden for c in range(3))
        W3[EP] = self._lap1(S2[EP], W1[EP], vt, self.S(sc["chipxyz"]))
        for c in range(3):
            W3[BX + c] = self._lap1(S2[BX + c], W1[BX + c], vt, self.chierg)
        W3 = b3(W3)
        # flux correction
        out = []
        for f in range(34):
            rec = None
            if record is not None:
                rec = []
                record.append(rec)
            chi = self.diferg if f >= BX else self.difrho
            out.append(self.fcsmooth(S[f], W3[f], chi,
                                     codes=None if codes is None else codes[f], record=rec))
        out = b3(out)
        return out + [W3[EP]]

    # --- nested boxes: code/sub_flanks.f90:5-303, code/sub_corer.f90:5-55 ---
    def flanks(self, X, flank, t1, t2):
        """The flanks on the 34 fields of X: per field, work = old t1 +
        cur t2 of the parent, its trilinear interpolation I0 at each node,
        and each face's two layers blended w = frac w + (1 - frac) I0
        (frac 0.5 on the inner layer, 0 on the outer one), the panels in the
        reference's order -- bottom, top, front, back, left, right."""
        g = {k: v if k in "ijk" else v.to(self.dt) for k, v in flank["geom"].items()}
        kb, jb, ib = g["k"][:, None, None], g["j"][None, :, None], g["i"][None, None, :]
        dz, ddz = g["dz"][:, None, None], g["ddz"][:, None, None]
        dy, ddy = g["dy"][None, :, None], g["ddy"][None, :, None]
        dx, ddx = g["dx"][None, None, :], g["ddx"][None, None, :]
        nz, ny, nx = X[0].shape
        half, zero, one = self.S(0.5), self.S(0.0), self.S(1.0)
        out = []
        for f in range(34):
            w = flank["old"][f] * t1 + flank["cur"][f] * t2
            I0 = (w[kb, jb, ib] * ddx * ddy * ddz
                  + w[kb + 1, jb, ib] * ddx * ddy * dz
                  + w[kb, jb + 1, ib] * ddx * dy * ddz
                  + w[kb + 1, jb + 1, ib] * ddx * dy * dz
                  + w[kb, jb, ib + 1] * dx * ddy * ddz
                  + w[kb + 1, jb, ib + 1] * dx * ddy * dz
                  + w[kb, jb + 1, ib + 1] * dx * dy * ddz
                  + w[kb + 1, jb + 1, ib + 1] * dx * dy * dz)
            a = X[f].clone()
            for axis, n in ((0, nz), (1, ny), (2, nx)):
                for lay in ((1, 0), (n - 2, n - 1)):
                    for idx, frac in zip(lay, (half, zero)):
                        sel = [slice(None)] * 3
                        sel[axis] = idx
                        sel = tuple(sel)
                        a[sel] = frac * a[sel] + (one - frac) * I0[sel]
            out.append(a)
        return out

    def corer(self, fine, coarse, geom):
        """coarse (the parent's state) with the core of its 34 fields
        replaced by the average of fine's (6/12 the centre, 1/12 each of its
        six neighbours) at every second fine node."""
        K, J, I = geom["fine"]
        KB, JB, IB = geom["coarse"]
        out = list(coarse)
        twelve = self.S(12.0)
        for f in range(34):
            q = fine[f]
            v = (self.six * q[K, J, I] + q[K, J, I + 1] + q[K, J, I - 1] + q[K, J + 1, I]
                 + q[K, J - 1, I] + q[K + 1, J, I] + q[K - 1, J, I]) / twelve
            a = coarse[f].clone()
            a[KB, JB, IB] = v
            out[f] = a
        return out


    def sample_full(self, st, plan, idx):
        return self.sample([st[i] for i in idx], plan)

    def rollout_full(self, ctl, st, steps, every, plan, idx):
        out = []
        for t in range(steps):
            st = self.step(st, ctl)
            if (t + 1) % every == 0:
                out.append(self.sample_full(st, plan, idx))
        return torch.stack(out)


class NestedRef:
    """The restatement of a nested stack (innermost first, as
    g3m3_amd.adjoint_nested.nested_setup builds it) on the full step's
    schedule (nested_schedule): each box a FullRef, inner boxes bounded by
    FullRef.flanks from their parent and restricted into it by
    FullRef.corer."""

    def __init__(self, setups, dtype=torch.float64):
        from g3m3_amd.adjoint_nested import nested_schedule
        self.setups = list(setups)
        self.n = len(self.setups)
        self.refs = [FullRef(s, dtype) for s in self.setups]
        self.fg = [flank_geom(self.setups[b], self.setups[b + 1]) for b in range(self.n - 1)]
        self.cg = [corer_geom(self.setups[b], self.setups[b + 1]) for b in range(self.n - 1)]
        self.events = nested_schedule([s.scal["delt"] for s in self.setups])

    def states(self, sts=None):
        sts = [s.state for s in self.setups] if sts is None else sts
        return [r.state(st) for r, st in zip(self.refs, sts)]

    def step(self, b, st, ctl, p_old=None, p_cur=None, t12=None, codes=None, record=None):
        fl = None if t12 is None else dict(old=p_old, cur=p_cur, t12=t12, geom=self.fg[b])
        return self.refs[b].step(st, ctl, codes=codes, record=record, flank=fl)

    def corer(self, b, child, parent):
        return self.refs[b].corer(child, parent, self.cg[b])

    def full_step(self, cur, ctl):
        cur, old = list(cur), [None] * self.n
        for ev in self.events:
            if ev[0] == "step":
                _, b, t12 = ev
                out = self.step(b, cur[b], ctl, *((old[b + 1], cur[b + 1], t12)
                                                  if t12 is not None else ()))
                old[b], cur[b] = cur[b], out
            else:
                b = ev[1]
                cur[b + 1] = self.corer(b, cur[b], cur[b + 1])
        return cur


def flank_geom(setup, parent):
    """The flanks' per-axis parent indices and weights of an inner box
    (code/sub_flanks.f90:24-30, 50-55), as torch tensors."""
    from oracle.restate import _axis_weights
    g = {}
    for ax, key, n in (("x", "i", setup.nx), ("y", "j", setup.ny), ("z", "k", setup.nz)):
        lo, d, dd = _axis_weights(setup.scal[ax + "min"], setup.scal[ax + "max"],
                                  parent.scal[ax + "min"], parent.scal[ax + "max"], n)
        g[key] = torch.from_numpy(lo)
        g["d" + ax] = torch.from_numpy(np.asarray(d, np.float64))
        g["dd" + ax] = torch.from_numpy(np.asarray(dd, np.float64))
    return g


def corer_geom(setup, parent):
    """The corer's fine and parent index grids (code/sub_corer.f90:17-32)."""
    F = np.float32
    sx = (F(setup.scal["xmax"]) - F(setup.scal["xmin"])) / F(setup.nx - 1.)
    d = {ax: (F(parent.scal[ax + "max"]) - F(parent.scal[ax + "min"])) / F(n - 1.)
         for ax, n in (("x", setup.nx), ("y", setup.ny), ("z", setup.nz))}
    nskip = int(d["x"] / sx)
    fine, coarse = [], []
    for ax, n in (("z", setup.nz), ("y", setup.ny), ("x", setup.nx)):
        nb1 = int(F(1.) + (F(setup.scal[ax + "min"]) - F(parent.scal[ax + "min"])) / d[ax]
                  + F(.5))
        s = np.arange(1 + nskip, n - nskip + 1, nskip)
        fine.append(s - 1)
        coarse.append(nb1 + (s - 1) // nskip - 1)
    mesh = lambda a: tuple(torch.from_numpy(x) for x in np.ix_(*a))
    return {"fine": mesh(fine), "coarse": mesh(coarse)}


def oracle_run(setup, ctl):
    """An OracleRun (oracle/boxstep.py) of the setup's box as a one-box
    run: mbndry = 0 (no planet lists), the setup's scalars and controls,
    wrkepres = the state's last array, the time step fixed at delt."""
    from g3m3_amd.workload import FIELDS, Workload, make_workload
    from oracle.boxstep import OracleRun

    nx, ny, nz = setup.nx, setup.ny, setup.nz
    base = make_workload("single61", dims=(nx, ny, nz), spacing=float(setup.scal["rx"]))
    sc = dict(base.scal)
    for k in ("rmassq", "rmassh", "rmasso", "reynolds", "smallbit", "gamma", "gamma1", "grav",
              "re_equiv", "chirho", "chipxyz", "chierg", "difrho", "diferg", "rho_frac",
              "o_conc", "ti_te"):
        sc[k] = np.float32(setup.scal[k])
    c = np.asarray(ctl, np.float32)
    sc.update(mbndry=0, tilting=False, eta=c[0], sbx_wind=c[1], sby_wind=c[2], sbz_wind=c[3],
              srho=c[4], spx=c[5], spy=c[6], spz=c[7], spress=c[8])
    shp = (1, nz, ny, nx)
    fields = {k: np.ascontiguousarray(setup.state[i].reshape(shp)) for i, k in enumerate(FIELDS)}
    for i, k in enumerate(("bx0", "by0", "bz0")):
        fields[k] = np.ascontiguousarray(setup.b0[i].reshape(shp))
    fields["rst"] = np.zeros(shp, np.float32)
    s = setup.scal
    grid = {"xspac": np.array([s["rx"]], np.float32)}
    for ax in "xyz":
        grid[f"grd_{ax}min"] = np.array([s[ax + "min"]], np.float32)
        grid[f"grd_{ax}max"] = np.array([s[ax + "max"]], np.float32)
    lists = {}
    for nm in ("zero", "mid", "srf"):
        lists["ij" + nm] = np.zeros((1, 3, 1), np.int32)
        lists["num" + nm] = np.zeros(1, np.int32)
        lists["parm_" + nm] = np.zeros((1, 7, 1), np.float32)
        lists["m" + nm] = 1
    wl = Workload(name="fullbox", nx=nx, ny=ny, nz=nz, ngrd=1, steps=1, divb_every=0,
                  divb_tol=1e-6, grid=grid, scal=sc, fields=fields, lists=lists)
    r = OracleRun(wl, "restate")
    r.a["wrkepres"][0] = setup.state[WE]
    return r


def nested_oracle_run(setups, ctl):
    """An OracleRun of a nested stack (innermost first, as
    g3m3_amd.adjoint_nested.nested_setup builds it): n_grids boxes with
    each setup's state, b0, wrkepres and grid, mbndry = 0, the controls of
    the outermost box."""
    from g3m3_amd.workload import FIELDS, Workload
    from oracle.boxstep import OracleRun

    one = oracle_run(setups[-1], ctl)
    n = len(setups)
    nx, ny, nz = setups[0].nx, setups[0].ny, setups[0].nz
    fields = {k: np.ascontiguousarray(np.stack([st.state[i] for st in setups]))
              for i, k in enumerate(FIELDS)}
    for i, k in enumerate(("bx0", "by0", "bz0")):
        fields[k] = np.ascontiguousarray(np.stack([st.b0[i] for st in setups]))
    fields["rst"] = np.zeros((n, nz, ny, nx), np.float32)
    grid = {"xspac": np.array([st.scal["rx"] for st in setups], np.float32)}
    for ax in "xyz":
        grid[f"grd_{ax}min"] = np.array([st.scal[ax + "min"] for st in setups], np.float32)
        grid[f"grd_{ax}max"] = np.array([st.scal[ax + "max"] for st in setups], np.float32)
    lists = {}
    for nm in ("zero", "mid", "srf"):
        lists["ij" + nm] = np.zeros((n, 3, 1), np.int32)
        lists["num" + nm] = np.zeros(n, np.int32)
        lists["parm_" + nm] = np.zeros((n, 7, 1), np.float32)
        lists["m" + nm] = 1
    wl = Workload(name="fullnest", nx=nx, ny=ny, nz=nz, ngrd=n, steps=1, divb_every=0,
                  divb_tol=1e-6, grid=grid, scal=dict(one.wl.scal), fields=fields,
                  lists=lists)
    r = OracleRun(wl, "restate")
    for b, st in enumerate(setups):
        r.a["wrkepres"][b] = st.state[WE]
    return r


def nested_oracle_full_step(r, setups):
    """One full step of the nested OracleRun with every box's time step its
    setup's delt; returns the boxes' new 35-array states."""
    from g3m3_amd.workload import FIELDS
    r.t_stepnew[:] = [float(np.float32(st.scal["delt"])) for st in setups]
    r.full_step()
    return [np.stack([r.a[k][b] for k in FIELDS] + [r.a["wrkepres"][b]])
            for b in range(len(setups))]


def oracle_step(r, delt):
    """One box-step of the one-box OracleRun with time step delt; returns
    the new 35-array state."""
    from g3m3_amd.workload import FIELDS
    r.t_step[0] = float(np.float32(delt))
    r.box_step(1)
    return np.stack([r.a[k][0] for k in FIELDS] + [r.a["wrkepres"][0]])


class Float32Values(TorchFunctionMode):
    """Within this mode every float64 tensor a torch function returns is
    rounded to the nearest float32 value, with an identity gradient.  For
    the operations of the restatement (+, -, *, /, sqrt, abs, min/max,
    copies), a float64 operation on float32 values rounded once to float32
    is the float32 operation, so a float64 run inside the mode takes the
    float32 run's values at every intermediate, and its backward pass, in
    float64, is the exact linearisation of the float32 step -- the
    function the device's float32 adjoint differentiates.  (Outside the
    mode the float64 backward linearises at float64 intermediates, which
    differ from float32 ones by up to ~1e-3 relative where the step
    differences nearly equal values: pressure gradients, velocity jumps.)"""

    def __torch_function__(self, func, types, args=(), kwargs=None):
        out = func(*args, **(kwargs or {}))
        if isinstance(out, torch.Tensor) and out.dtype == torch.float64:
            with torch.no_grad():
                r = out.to(torch.float32).to(torch.float64)
            out = out + (r - out).detach() if out.requires_grad else r
        return out
