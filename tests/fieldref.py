"""Test infrastructure: a torch restatement of the frozen-plasma field step
(g3m3_field_model_step) and of craft sampling, written for torch.autograd.

Run in float64 it gives the reference gradient the device adjoint is held
to; run in float32 it must reproduce the oracle's composition of the same
reference subroutines (tests/fieldoracle.py), which pins it.  fcsmooth's
limiter is not differentiable where two of its arguments tie; the float32
run can record its branches and a float64 run can take them, which gives
the float64 linearisation of the float32 model the device differentiates.  Arrays are
(nz, ny, nx); b is a tuple of three such tensors.

Reference subroutines followed: calcur code/sub_calcur.f90:5-189,
fnd_evel code/sub_fnd_evel.f90:4-57, BandE code/sub_BandE.f90:5-162,
push_bfld code/sub_push_bfld.f90:4-50, bndry_outer
code/sub_bndry_outer.f90:50-728 (field rules), lap_bfld
code/sub_lap_bfld.f90:4-67, fcsmooth code/sub_fcsmooth.f90:7-242,
findgrid code/cft_findgrid.f90:4-60, trilin_interp
code/cft_trilin_interp.f90:10-66.
"""
from __future__ import annotations

import numpy as np
import torch


def _set(a, idx, v):
    a = a.clone()
    a[idx] = v
    return a


class FieldRef:
    def __init__(self, setup, dtype=torch.float64):
        self.dt = dtype
        T = lambda a: torch.from_numpy(np.asarray(a, np.float32)).to(dtype)
        S = lambda x: torch.tensor(float(np.float32(x)), dtype=dtype)
        self.s = setup
        pl = {k: T(v) for k, v in setup.plasma.items()}
        self.pl = pl
        self.b0 = tuple(T(setup.b0[c]) for c in range(3))
        sc = setup.scal
        self.rx, self.delt = S(sc["rx"]), S(sc["delt"])
        self.rmq, self.rmh, self.rmo = S(sc["rmassq"]), S(sc["rmassh"]), S(sc["rmasso"])
        self.re, self.chierg, self.diferg = S(sc["reynolds"]), S(sc["chierg"]), S(sc["diferg"])
        one = S(1.0)
        self.two, self.three, self.six, self.half = S(2.0), S(3.0), S(6.0), S(0.5)
        d = S(0.000001)
        q, h, o = pl["qrho"], pl["hrho"], pl["orho"]
        self.tot = (q + d) / self.rmq + (h + d) / self.rmh + (o + d) / self.rmo
        tden = q + h + o
        self.tv = tuple((pl["q" + c] + pl["h" + c] + pl["o" + c]) / tden
                        for c in ("px", "py", "pz"))
        self.eden = q / self.rmq + h / self.rmh + o / self.rmo
        self.arho = self.eden * self.re
        del one

    # interior views of a (nz, ny, nx) tensor, shifted by (dz, dy, dx)
    @staticmethod
    def _i(a, dz=0, dy=0, dx=0):
        nz, ny, nx = a.shape
        return a[1 + dz:nz - 1 + dz, 1 + dy:ny - 1 + dy, 1 + dx:nx - 1 + dx]

    def calcur(self, b):
        I = self._i
        X, Y, Z = b
        dxt, dyt, dzt = self.two * self.rx, self.two * self.rx, self.two * self.rx
        cx = torch.zeros_like(X)
        cy = torch.zeros_like(X)
        cz = torch.zeros_like(X)
        cx = _set(cx, (slice(1, -1),) * 3, (I(Z, 0, 1) - I(Z, 0, -1)) / dyt - (I(Y, 1) - I(Y, -1)) / dzt)
        cy = _set(cy, (slice(1, -1),) * 3, (I(X, 1) - I(X, -1)) / dzt - (I(Z, 0, 0, 1) - I(Z, 0, 0, -1)) / dxt)
        cz = _set(cz, (slice(1, -1),) * 3, (I(Y, 0, 0, 1) - I(Y, 0, 0, -1)) / dxt - (I(X, 0, 1) - I(X, 0, -1)) / dyt)
        out = []
        for c in (cx, cy, cz):
            c = c.clone()
            # panels: z faces, x faces, y faces (:58-104)
            c[0, 1:-1, 1:-1] = c[1, 1:-1, 1:-1]
            c[-1, 1:-1, 1:-1] = c[-2, 1:-1, 1:-1]
            c[1:-1, 1:-1, 0] = c[1:-1, 1:-1, 1]
            c[1:-1, 1:-1, -1] = c[1:-1, 1:-1, -2]
            c[1:-1, 0, 1:-1] = c[1:-1, 1, 1:-1]
            c[1:-1, -1, 1:-1] = c[1:-1, -2, 1:-1]
            # lines along x (:109-121)
            c[0, 0, 1:-1] = c[1, 0, 1:-1]
            c[-1, 0, 1:-1] = c[-2, 0, 1:-1]
            c[0, -1, 1:-1] = c[1, -1, 1:-1]
            c[-1, -1, 1:-1] = c[-1, -2, 1:-1]
            # lines along y (:124-138)
            c[0, 1:-1, 0] = c[1, 1:-1, 0]
            c[-1, 1:-1, 0] = c[-2, 1:-1, 0]
            c[0, 1:-1, -1] = c[1, 1:-1, -1]
            c[-1, 1:-1, -1] = c[-2, 1:-1, -1]
            # lines along z (:141-155)
            c[1:-1, 0, 0] = c[1:-1, 1, 0]
            c[1:-1, 0, -1] = c[1:-1, 1, -1]
            c[1:-1, -1, 0] = c[1:-1, -2, 0]
            c[1:-1, -1, -1] = c[1:-1, -2, -1]
            # corners (:157-186)
            c[0, 0, 0] = c[1, 0, 0]
            c[0, -1, 0] = c[1, -1, 0]
            c[-1, 0, 0] = (c[-2, 0, 0] + c[-1, 1, 0]) / self.two
            c[-1, -1, 0] = (c[-2, -1, 0] + c[-1, -2, 0]) / self.two
            c[0, 0, -1] = (c[1, 0, -1] + c[0, 1, -1] + c[0, 0, -2]) / self.three
            c[0, -1, -1] = (c[1, -1, -1] + c[0, -2, -1] + c[0, -1, -2]) / self.three
            c[-1, 0, -1] = (c[-2, 0, -1] + c[-1, 1, -1] + c[-1, 0, -2]) / self.three
            c[-1, -1, -1] = (c[-2, -1, -1] + c[-1, -2, -1] + c[-1, -1, -2]) / self.three
            out.append(c)
        return tuple(out)

    def efield(self, b, eta):
        I = self._i
        J = self.calcur(b)
        bs = tuple(self.b0[c] + b[c] for c in range(3))
        ev = tuple(self.tv[c] - J[c] / self.tot / self.re for c in range(3))
        ex = -(ev[1] * bs[2] - ev[2] * bs[1])
        ey = -(ev[2] * bs[0] - ev[0] * bs[2])
        ez = -(ev[0] * bs[1] - ev[1] * bs[0])
        dxt = self.two * self.rx
        ep = self.pl["epres"]
        ar = I(self.arho)
        ex = _set(ex, (slice(1, -1),) * 3, I(ex) - ((I(ep, 0, 0, 1) - I(ep, 0, 0, -1)) / dxt) / ar)
        ey = _set(ey, (slice(1, -1),) * 3, I(ey) - ((I(ep, 0, 1) - I(ep, 0, -1)) / dxt) / ar)
        ez = _set(ez, (slice(1, -1),) * 3, I(ez) - ((I(ep, 1) - I(ep, -1)) / dxt) / ar)
        # uniform resistivity on the interior, before the face copies
        en = I(self.eden)
        E = []
        for e, j in zip((ex, ey, ez), J):
            e = _set(e, (slice(1, -1),) * 3, I(e) + eta * I(j) / en)
            E.append(e)
        out = []
        for e in E:
            e = e.clone()
            e[:, :, 0] = e[:, :, 1]
            e[:, :, -1] = e[:, :, -2]
            out.append(e)
        E, out = out, []
        for e in E:
            e = e.clone()
            e[0] = e[1]
            e[-1] = e[-2]
            out.append(e)
        E, out = out, []
        for e in E:
            e = e.clone()
            e[:, 0, :] = e[:, 1, :]
            e[:, -1, :] = e[:, -2, :]
            out.append(e)
        return tuple(out)

    def push(self, old, E, delt):
        I = self._i
        dxt = self.two * self.rx
        ex, ey, ez = E
        nx = _set(old[0], (slice(1, -1),) * 3, I(old[0]) - delt * (
            ((I(ez, 0, 1) - I(ez, 0, -1)) / dxt) - ((I(ey, 1) - I(ey, -1)) / dxt)))
        ny = _set(old[1], (slice(1, -1),) * 3, I(old[1]) + delt * (
            ((I(ez, 0, 0, 1) - I(ez, 0, 0, -1)) / dxt) - ((I(ex, 1) - I(ex, -1)) / dxt)))
        nz = _set(old[2], (slice(1, -1),) * 3, I(old[2]) - delt * (
            ((I(ey, 0, 0, 1) - I(ey, 0, 0, -1)) / dxt) - ((I(ex, 0, 1) - I(ex, 0, -1)) / dxt)))
        return (nx, ny, nz)

    def outer(self, b, sb):
        out = []
        for comp in range(3):
            f = b[comp].clone()
            f[:, :, 0] = sb[comp] - self.b0[comp][:, :, 0]
            f[:, :, -1] = f[:, :, -2]
            f[:, 0, 1:] = f[:, 1, 1:]
            f[:, -1, 1:] = f[:, -2, 1:]
            f[-1, :, 1:] = f[-2, :, 1:]
            f[0, :, 1:] = f[1, :, 1:]
            if comp == 0:
                f[1:-1, 0, -1] = f[1:-1, 0, -2]
                f[1:-1, -1, -1] = f[1:-1, -1, -2]
            else:
                f[1:-1, 0, -1] = (f[1:-1, 0, -2] + f[1:-1, 1, -1]) / self.two
                f[1:-1, -1, -1] = (f[1:-1, -1, -2] + f[1:-1, -2, -1]) / self.two
            f[0, 1:-1, -1] = f[1, 1:-1, -1]
            f[-1, 1:-1, -1] = f[-1, 1:-1, -2]
            f[0, 0, 1:] = f[1, 0, 1:]
            f[0, -1, 1:] = f[1, -1, 1:]
            f[-1, 0, 1:] = f[-2, 0, 1:]
            f[-1, -1, 1:] = f[-2, -1, 1:]
            out.append(f)
        return tuple(out)

    def lap(self, q, base):
        I = self._i
        vx, vy, vz = self.tv_lap
        uxp = torch.abs(I(vx, 0, 0, 1) - I(vx))
        uxm = torch.abs(I(vx) - I(vx, 0, 0, -1))
        uyp = torch.abs(I(vy, 0, 1) - I(vy))
        uym = torch.abs(I(vy) - I(vy, 0, -1))
        uzp = torch.abs(I(vz, 1) - I(vz))
        uzm = torch.abs(I(vz) - I(vz, -1))
        out = []
        for qq, w in zip(q, base):
            c = I(qq)
            d = self.delt
            v = c + self.chierg * (d * (uxp * (I(qq, 0, 0, 1) - c) - uxm * (c - I(qq, 0, 0, -1)))
                                   + d * (uyp * (I(qq, 0, 1) - c) - uym * (c - I(qq, 0, -1)))
                                   + d * (uzp * (I(qq, 1) - c) - uzm * (c - I(qq, -1))))
            out.append(_set(w, (slice(1, -1),) * 3, v))
        return tuple(out)

    @property
    def tv_lap(self):
        # fnd_vtot of the fixed plasma equals fnd_evel's tv
        return self.tv

    def fcsmooth(self, O, W, chi, codes=None, record=None):
        """codes: the limiter's branch per axis (as recorded), used instead
        of this precision's own choice; record: a list the branches this
        evaluation takes are appended to.  Branch 1 = d1(f-1), 2 = chi d(f),
        3 = d1(f+1), 0 = zero flux; the first strict minimum wins and a
        minimum that is not positive gives 0 (the device's rule)."""
        I = self._i
        P = _set(W, (slice(1, -1),) * 3, I(W) + chi * (
            I(O, 0, 0, 1) + I(O, 0, 0, -1) + I(O, 0, 1) + I(O, 0, -1) + I(O, 1) + I(O, -1)
            - self.six * I(O)))

        def lim(s, d1m, chiabs, d1p):
            return s * torch.clamp(torch.minimum(torch.minimum(s * d1m, chiabs), s * d1p), min=0)

        def axis_flux(dim, ax):
            n = W.shape[dim]
            d = torch.diff(W, dim=dim)
            d1 = torch.diff(P, dim=dim)
            d = torch.cat([d, d.narrow(dim, n - 2, 1)], dim)
            d1 = torch.cat([d1, d1.narrow(dim, n - 2, 1)], dim)
            dm = d.narrow(dim, 1, n - 2)
            s = torch.sign(dm)
            s = torch.where(s == 0, torch.where(torch.signbit(dm), -1.0, 1.0).to(dm.dtype), s)
            s = s.detach()
            d1m, d1p = d1.narrow(dim, 0, n - 2), d1.narrow(dim, 2, n - 2)
            if codes is not None:
                br = codes[ax]
                z = torch.zeros_like(dm)
                fn = torch.where(br == 1, d1m, torch.where(br == 2, chi * dm,
                                                           torch.where(br == 3, d1p, z)))
            else:
                fn = lim(s, d1m, chi * torch.abs(dm), d1p)
            if record is not None:
                with torch.no_grad():
                    v1, v2, v3 = s * d1m, chi * torch.abs(dm), s * d1p
                    br = torch.ones(v1.shape, dtype=torch.int8)
                    mn = v1
                    m2 = v2 < mn
                    br = torch.where(m2, 2, br).to(torch.int8)
                    mn = torch.where(m2, v2, mn)
                    m3 = v3 < mn
                    br = torch.where(m3, 3, br).to(torch.int8)
                    mn = torch.where(m3, v3, mn)
                    record.append(torch.where(mn > 0, br, 0).to(torch.int8))
            return torch.cat([fn.narrow(dim, 0, 1), fn, fn.narrow(dim, n - 3, 1)], dim)

        fx = axis_flux(2, 0)
        T = W.clone()
        T = _set(T, (slice(None), slice(None), slice(1, -1)),
                 P[:, :, 1:-1] - fx[:, :, 1:-1] + fx[:, :, :-2])
        fy = axis_flux(1, 1)
        T = _set(T, (slice(None), slice(1, -1), slice(None)),
                 T[:, 1:-1, :] - fy[:, 1:-1, :] + fy[:, :-2, :])
        fz = axis_flux(0, 2)
        out = _set(P, (slice(1, -1),), T[1:-1] - fz[1:-1] + fz[:-2])
        return out

    def step(self, b, ctl, codes=None, record=None):
        """One step.  codes / record: fcsmooth's limiter branches per
        component (lists of three per-axis tensors), see fcsmooth."""
        eta, sb = ctl[0], (ctl[1], ctl[2], ctl[3])
        E1 = self.efield(b, eta)
        w1 = self.outer(self.push(b, E1, self.half * self.delt), sb)
        E2 = self.efield(w1, eta)
        b2 = self.outer(self.push(b, E2, self.delt), sb)
        W = self.outer(self.lap(b2, w1), sb)
        bn = []
        for c in range(3):
            rec = None
            if record is not None:
                rec = []
                record.append(rec)
            bn.append(self.fcsmooth(b[c], W[c], self.diferg,
                                    codes=None if codes is None else codes[c], record=rec))
        return self.outer(tuple(bn), sb)

    # --- craft sampling: findgrid + aux cube + trilin_interp on one box ---
    def sample_plan(self, pos):
        """Per point: the cube's lower corner (0-based i, j, k) and the
        fractions (xd, yd, zd), computed in float32 as the device does."""
        g = self.s.grid
        F = np.float32
        mn, xs, re = g["grid_minvals"][0], F(g["xspac"][0]), F(g["re_equiv"])
        plan = []
        for p in np.asarray(pos, F):
            xa = (p / re).astype(F)
            gp = [1 + int(F(F(xa[a] - mn[a]) / xs)) for a in range(3)]
            v0 = [F((mn[a] + (F(gp[a]) - F(1)) * xs) * re) for a in range(3)]
            v1 = [F(v0[a] + xs * re) for a in range(3)]
            fr = [F(F(p[a] - v0[a]) / F(v1[a] - v0[a])) for a in range(3)]
            plan.append(([gp[0] - 1, gp[1] - 1, gp[2] - 1], fr))
        return plan

    def sample(self, b, plan):
        rows = []
        one = torch.tensor(1.0, dtype=self.dt)
        for (i, j, k), fr in plan:
            xd, yd, zd = (torch.tensor(float(f), dtype=self.dt) for f in fr)
            vals = []
            for q in b:
                c = q[k:k + 2, j:j + 2, i:i + 2]
                e00 = c[0, 0, 0] * (one - xd) + c[0, 0, 1] * xd
                e01 = c[1, 0, 0] * (one - xd) + c[1, 0, 1] * xd
                e10 = c[0, 1, 0] * (one - xd) + c[0, 1, 1] * xd
                e11 = c[1, 1, 0] * (one - xd) + c[1, 1, 1] * xd
                f0 = e00 * (one - yd) + e10 * yd
                f1 = e01 * (one - yd) + e11 * yd
                vals.append(f0 * (one - zd) + f1 * zd)
            rows.append(torch.stack(vals))
        return torch.stack(rows)

    def rollout(self, ctl, b, steps, every, plan):
        out = []
        for t in range(steps):
            b = self.step(b, ctl)
            if (t + 1) % every == 0:
                out.append(self.sample(b, plan))
        return torch.stack(out)

    def state(self, b_np):
        return tuple(torch.from_numpy(np.asarray(b_np[c], np.float32)).to(self.dt)
                     for c in range(3))
