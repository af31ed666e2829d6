"""NumPy restatement of G3M^3's per-timestep hot-path subroutines.

TEST INFRASTRUCTURE ONLY -- this is the oracle.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may
import it, and only as the checker (or as the CPU-port baseline when the
reference build is absent).  The product (``g3m3_amd``) never imports it.

Every function keeps the reference subroutine's name and positional
argument list, so a test can hand one argument list to the reference build
(``oracle.ref``) and to this module.  Conventions:

* a field ``q(nx,ny,nz,n_grids)`` is a float32 numpy array of shape
  ``(n_grids, nz, ny, nx)`` (C order, byte-identical to the Fortran array);
  a box-local scratch ``q(nx,ny,nz)`` has shape ``(nz, ny, nx)``;
  ``box`` is 1-based, as in the reference;
* arrays are updated in place, as the Fortran dummies are;
* arithmetic is IEEE single precision, with each expression evaluated in
  the reference's left-to-right order and with no fused multiply-add, so
  results match a serial ``-ffp-contract=off`` build of the reference bit
  for bit (the pinning tests check exactly that);
* scalar outputs (``set_speed_agrd``'s maxima, ``sparse_solver``'s
  ``residual2``, ``trilin_interp``'s ``scdata``) are 1-element arrays.

Parity status per function is recorded in the test files that pin it
(tests/test_oracle_vs_ref.py, tests/test_oracle_invariants.py).
"""
from __future__ import annotations

import numpy as np

F = np.float32
I32 = np.int32

# ---------------------------------------------------------------------------
# stencil views of the interior 2..n-1 (1-based) of a (nz, ny, nx) array


def _sh(dz, dy, dx):
    return (slice(1 + dz, (-1 + dz) or None),
            slice(1 + dy, (-1 + dy) or None),
            slice(1 + dx, (-1 + dx) or None))


C = _sh(0, 0, 0)
XP, XM = _sh(0, 0, 1), _sh(0, 0, -1)
YP, YM = _sh(0, 1, 0), _sh(0, -1, 0)
ZP, ZM = _sh(1, 0, 0), _sh(-1, 0, 0)


def _f(*xs):
    return [F(x) for x in xs]


# ---------------------------------------------------------------------------
# code/sub_store_array.f90:1-20


def store_array(bx, bx0, nx, ny, nz, n_grids, box):
    bx[box - 1][...] = bx0[box - 1]


# ---------------------------------------------------------------------------
# code/sub_calcur.f90:5-189  J = curl(b), face/edge/corner rules


def calcur(bx, by, bz, nx, ny, nz, n_grids, box, curx, cury, curz, rx, ry, rz):
    b = box - 1
    rx, ry, rz = _f(rx, ry, rz)
    dxt, dyt, dzt = F(2) * rx, F(2) * ry, F(2) * rz
    X, Y, Z = bx[b], by[b], bz[b]
    # interior (:36-44)
    curx[C] = (Z[YP] - Z[YM]) / dyt - (Y[ZP] - Y[ZM]) / dzt
    cury[C] = (X[ZP] - X[ZM]) / dzt - (Z[XP] - Z[XM]) / dxt
    curz[C] = (Y[XP] - Y[XM]) / dxt - (X[YP] - X[YM]) / dyt
    for c in (curx, cury, curz):
        # bottom/top panels (:58-72)
        c[0, 1:-1, 1:-1] = c[1, 1:-1, 1:-1]
        c[-1, 1:-1, 1:-1] = c[-2, 1:-1, 1:-1]
        # front/back (:76-88)
        c[1:-1, 1:-1, 0] = c[1:-1, 1:-1, 1]
        c[1:-1, 1:-1, -1] = c[1:-1, 1:-1, -2]
        # left/right (:92-104)
        c[1:-1, 0, 1:-1] = c[1:-1, 1, 1:-1]
        c[1:-1, -1, 1:-1] = c[1:-1, -2, 1:-1]
        # corner lines along x (:109-121); note (i,ny,nz) <- (i,ny1,nz)
        c[0, 0, 1:-1] = c[1, 0, 1:-1]
        c[-1, 0, 1:-1] = c[-2, 0, 1:-1]
        c[0, -1, 1:-1] = c[1, -1, 1:-1]
        c[-1, -1, 1:-1] = c[-1, -2, 1:-1]
        # corner lines along y (:124-138)
        c[0, 1:-1, 0] = c[1, 1:-1, 0]
        c[-1, 1:-1, 0] = c[-2, 1:-1, 0]
        c[0, 1:-1, -1] = c[1, 1:-1, -1]
        c[-1, 1:-1, -1] = c[-2, 1:-1, -1]
        # corner lines along z (:141-155)
        c[1:-1, 0, 0] = c[1:-1, 1, 0]
        c[1:-1, 0, -1] = c[1:-1, 1, -1]
        c[1:-1, -1, 0] = c[1:-1, -2, 0]
        c[1:-1, -1, -1] = c[1:-1, -2, -1]
        # corner points (:157-186)
        c[0, 0, 0] = c[1, 0, 0]
        c[0, -1, 0] = c[1, -1, 0]
        c[-1, 0, 0] = (c[-2, 0, 0] + c[-1, 1, 0]) / F(2)
        c[-1, -1, 0] = (c[-2, -1, 0] + c[-1, -2, 0]) / F(2)
        c[0, 0, -1] = (c[1, 0, -1] + c[0, 1, -1] + c[0, 0, -2]) / F(3)
        c[0, -1, -1] = (c[1, -1, -1] + c[0, -2, -1] + c[0, -1, -2]) / F(3)
        c[-1, 0, -1] = (c[-2, 0, -1] + c[-1, 1, -1] + c[-1, 0, -2]) / F(3)
        c[-1, -1, -1] = (c[-2, -1, -1] + c[-1, -2, -1] + c[-1, -1, -2]) / F(3)


# ---------------------------------------------------------------------------
# code/sub_fnd_evel.f90:4-57


def fnd_evel(qpx, qpy, qpz, qrho, hpx, hpy, hpz, hrho, opx, opy, opz, orho,
             curx, cury, curz, evx, evy, evz, tvx, tvy, tvz,
             nx, ny, nz, n_grids, box, rmassq, rmassh, rmasso, reynolds):
    b = box - 1
    rmassq, rmassh, rmasso, reynolds = _f(rmassq, rmassh, rmasso, reynolds)
    d = F(0.000001)
    qden = (qrho[b] + d) / rmassq
    hden = (hrho[b] + d) / rmassh
    oden = (orho[b] + d) / rmasso
    tot = qden + hden + oden
    tden = qrho[b] + hrho[b] + orho[b]
    tvx[...] = (qpx[b] + hpx[b] + opx[b]) / tden
    evx[...] = tvx - curx / tot / reynolds
    tvy[...] = (qpy[b] + hpy[b] + opy[b]) / tden
    evy[...] = tvy - cury / tot / reynolds
    tvz[...] = (qpz[b] + hpz[b] + opz[b]) / tden
    evz[...] = tvz - curz / tot / reynolds


# ---------------------------------------------------------------------------
# code/sub_BandE.f90:5-162  Ohm's law


def BandE(efldx, efldy, efldz, bsx, bsy, bsz, curx, cury, curz, evx, evy, evz,
          btot, epres, qrho, hrho, orho, rst, resist, reynolds,
          nx, ny, nz, n_grids, box, rmassq, rmassh, rmasso,
          ijmid, nummid, ijzero, mbndry, numzero, mmid, mzero, rx, ry, rz):
    b = box - 1
    rmassq, rmassh, rmasso, reynolds, resist = _f(rmassq, rmassh, rmasso, reynolds, resist)
    rx, ry, rz = _f(rx, ry, rz)
    # -(ve x B) everywhere (:32-51)
    efldx[...] = -(evy * bsz - evz * bsy)
    efldy[...] = -(evz * bsx - evx * bsz)
    efldz[...] = -(evx * bsy - evy * bsx)
    # grad pe on the interior (:55-84)
    dxt, dyt, dzt = F(2) * rx, F(2) * ry, F(2) * rz
    q, h, o, ep = qrho[b], hrho[b], orho[b], epres[b]
    arho = (q[C] / rmassq + h[C] / rmassh + o[C] / rmasso) * reynolds
    efldx[C] = efldx[C] - ((ep[XP] - ep[XM]) / dxt) / arho
    efldy[C] = efldy[C] - ((ep[YP] - ep[YM]) / dyt) / arho
    efldz[C] = efldz[C] - ((ep[ZP] - ep[ZM]) / dzt) / arho
    # resistive term on the inner-boundary lists (:88-116)
    if box <= mbndry and resist < F(5000.):
        for lst, num in ((ijmid, nummid), (ijzero, numzero)):
            n = int(num[b])
            if n == 0:
                continue
            ii = lst[:n, 0, b] - 1
            jj = lst[:n, 1, b] - 1
            kk = lst[:n, 2, b] - 1
            eden = (q[kk, jj, ii] / rmassq + h[kk, jj, ii] / rmassh
                    + o[kk, jj, ii] / rmasso)
            r = rst[b][kk, jj, ii]
            efldx[kk, jj, ii] = efldx[kk, jj, ii] + r * curx[kk, jj, ii] / eden
            efldy[kk, jj, ii] = efldy[kk, jj, ii] + r * cury[kk, jj, ii] / eden
            efldz[kk, jj, ii] = efldz[kk, jj, ii] + r * curz[kk, jj, ii] / eden
    # flank copies, x then z then y (:125-159)
    for e in (efldx, efldy, efldz):
        e[:, :, 0] = e[:, :, 1]
        e[:, :, -1] = e[:, :, -2]
    for e in (efldx, efldy, efldz):
        e[0] = e[1]
        e[-1] = e[-2]
    for e in (efldx, efldy, efldz):
        e[:, 0, :] = e[:, 1, :]
        e[:, -1, :] = e[:, -2, :]


# ---------------------------------------------------------------------------
# code/sub_push_elec.f90:4-53


def push_elec(wrkepres, oldepres, epres, evx, evy, evz, gamma, gamma1,
              nx, ny, nz, n_grids, box, delt, rx, ry, rz):
    b = box - 1
    gamma, gamma1, delt, rx, ry, rz = _f(gamma, gamma1, delt, rx, ry, rz)
    dxt, dyt, dzt = F(2) * rx, F(2) * ry, F(2) * rz
    w = wrkepres[b]
    egx = (w[XP] - w[XM]) / dxt
    egy = (w[YP] - w[YM]) / dyt
    egz = (w[ZP] - w[ZM]) / dzt
    e = oldepres[b][C] - delt * gamma * (
        ((w[XP] * evx[XP] - w[XM] * evx[XM]) / dxt)
        + ((w[YP] * evy[YP] - w[YM] * evy[YM]) / dyt)
        + ((w[ZP] * evz[ZP] - w[ZM] * evz[ZM]) / dzt))
    epres[b][C] = e + delt * gamma1 * (evx[C] * egx + evy[C] * egy + evz[C] * egz)


# ---------------------------------------------------------------------------
# code/sub_push_bfld.f90:4-50  induction


def push_bfld(bx, by, bz, oldbx, oldby, oldbz, efldx, efldy, efldz,
              nx, ny, nz, n_grids, box, delt, rx, ry, rz):
    b = box - 1
    delt, rx, ry, rz = _f(delt, rx, ry, rz)
    dxt, dyt, dzt = F(2) * rx, F(2) * ry, F(2) * rz
    ex, ey, ez = efldx, efldy, efldz
    bx[b][C] = oldbx[b][C] - delt * (((ez[YP] - ez[YM]) / dyt) - ((ey[ZP] - ey[ZM]) / dzt))
    by[b][C] = oldby[b][C] + delt * (((ez[XP] - ez[XM]) / dxt) - ((ex[ZP] - ex[ZM]) / dzt))
    bz[b][C] = oldbz[b][C] - delt * (((ey[XP] - ey[XM]) / dxt) - ((ex[YP] - ex[YM]) / dyt))


# ---------------------------------------------------------------------------
# code/sub_push_ion.f90:4-362  two-step Lax-Wendroff ion push, one species


def push_ion(qrho, qpresx, qpresy, qpresz, qpx, qpy, qpz,
             oldqrho, oldqpresx, oldqpresy, oldqpresz, oldqpx, oldqpy, oldqpz,
             wrkqrho, wrkqpresx, wrkqpresy, wrkqpresz, wrkqpx, wrkqpy, wrkqpz,
             qpresxy, qpresxz, qpresyz, oldqpresxy, oldqpresxz, oldqpresyz,
             wrkqpresxy, wrkqpresxz, wrkqpresyz,
             bsx, bsy, bsz, btot, efldx, efldy, efldz, rmassq,
             qvx, qvy, qvz, tvx, tvy, tvz, gamma, gamma1,
             nx, ny, nz, n_grids, box, delt, grav, re_equiv, reynolds,
             grd_xmin, grd_xmax, grd_ymin, grd_ymax, grd_zmin, grd_zmax, isotropic):
    b = box - 1
    rmassq, delt, grav, re_equiv, reynolds = _f(rmassq, delt, grav, re_equiv, reynolds)
    ddx = (F(grd_xmax[b]) - F(grd_xmin[b])) / F(nx - 1)
    ddy = (F(grd_ymax[b]) - F(grd_ymin[b])) / F(ny - 1)
    ddz = (F(grd_zmax[b]) - F(grd_zmin[b])) / F(nz - 1)
    dxt, dyt, dzt = F(2) * ddx, F(2) * ddy, F(2) * ddz
    d_min = F(1e-6)
    rmass = rmassq
    bmax = F(0.20) * rmassq / delt / reynolds

    wr, wpx, wpy, wpz = wrkqrho[b], wrkqpx[b], wrkqpy[b], wrkqpz[b]
    wx, wy, wz = wrkqpresx[b], wrkqpresy[b], wrkqpresz[b]
    wxy, wxz, wyz = wrkqpresxy[b], wrkqpresxz[b], wrkqpresyz[b]

    # velocities (:67-81)
    aq = np.maximum(wr, d_min)
    qvx[...] = wpx / aq
    qvy[...] = wpy / aq
    qvz[...] = wpz / aq

    def flux(w):
        return (((w[XP] * qvx[XP] - w[XM] * qvx[XM]) / dxt)
                + ((w[YP] * qvy[YP] - w[YM] * qvy[YM]) / dyt)
                + ((w[ZP] * qvz[ZP] - w[ZM] * qvz[ZM]) / dzt))

    # continuity (:107-110)
    qrho[b][C] = oldqrho[b][C] - delt * (((wpx[XP] - wpx[XM]) / dxt)
                                        + ((wpy[YP] - wpy[YM]) / dyt)
                                        + ((wpz[ZP] - wpz[ZM]) / dzt))
    aqpx = oldqpx[b][C] - delt * flux(wpx)
    aqpy = oldqpy[b][C] - delt * flux(wpy)
    aqpz = oldqpz[b][C] - delt * flux(wpz)

    abx, aby, abz = bsx[C], bsy[C], bsz[C]
    bmag = btot[C]
    weight = np.maximum(F(1), bmag / bmax)
    skin_depth = reynolds / weight

    tx, ty, tz = tvx[C], tvy[C], tvz[C]
    vx, vy, vz = qvx[C], qvy[C], qvz[C]
    tcx = (ty * abz - tz * aby)
    tcy = -(tx * abz - tz * abx)
    tcz = (tx * aby - ty * abx)
    qcx = (vy * abz - vz * aby)
    qcy = -(vx * abz - vz * abx)
    qcz = (vx * aby - vy * abx)

    gpx = (wx[XP] - wx[XM]) / dxt + (wxy[YP] - wxy[YM]) / dxt + (wxz[ZP] - wxz[ZM]) / dxt
    gpy = (wxy[XP] - wxy[XM]) / dxt + (wy[YP] - wy[YM]) / dyt + (wyz[ZP] - wyz[ZM]) / dxt
    gpz = (wxz[XP] - wxz[XM]) / dxt + (wyz[YP] - wyz[YM]) / dyt + (wz[ZP] - wz[ZM]) / dzt

    dex = (efldx[C] + tcx) * reynolds
    dey = (efldy[C] + tcy) * reynolds
    dez = (efldz[C] + tcz) * reynolds
    dvx = (qcx - tcx) * skin_depth
    dvy = (qcy - tcy) * skin_depth
    dvz = (qcz - tcz) * skin_depth

    qden = wr[C] / rmassq
    px_ = aqpx + delt * (qden * (dex + dvx) - gpx)
    py_ = aqpy + delt * (qden * (dey + dvy) - gpy)
    pz_ = aqpz + delt * (qden * (dez + dvz) - gpz)

    # gravity (:202-206)
    ax = (F(grd_xmin[b]) + ddx * np.arange(1, nx - 1, dtype=F))[None, None, :]
    ay = (F(grd_ymin[b]) + ddy * np.arange(1, ny - 1, dtype=F))[None, :, None]
    az = (F(grd_zmin[b]) + ddz * np.arange(1, nz - 1, dtype=F))[:, None, None]
    radius = np.sqrt(ax * ax + ay * ay + az * az) + F(0.0000001)
    t = re_equiv * radius
    g = delt * wr[C] * grav / (t * t)
    qpx[b][C] = px_ - g * ax / radius
    qpy[b][C] = py_ - g * ay / radius
    qpz[b][C] = pz_ - g * az / radius

    two = F(2)
    # diagonal pressures, isotropic part (:210-238)
    px0 = oldqpresx[b][C] - delt * flux(wx) - two * delt * wx[C] * (qvx[XP] - qvx[XM]) / dxt
    py0 = oldqpresy[b][C] - delt * flux(wy) - two * delt * wy[C] * (qvy[YP] - qvy[YM]) / dyt
    pz0 = oldqpresz[b][C] - delt * flux(wz) - two * delt * wz[C] * (qvz[ZP] - qvz[ZM]) / dzt

    # smoothed B for the gyrotropic terms (:242-254)
    abx = (F(6) * bsx[C] + bsx[XP] + bsx[XM] + bsx[YP] + bsx[YM] + bsx[ZP] + bsx[ZM]) / F(12)
    aby = (F(6) * bsy[C] + bsy[XP] + bsy[XM] + bsy[YP] + bsy[YM] + bsy[ZP] + bsy[ZM]) / F(12)
    abz = (F(6) * bsz[C] + bsz[XP] + bsz[XM] + bsz[YP] + bsz[YM] + bsz[ZP] + bsz[ZM]) / F(12)
    bmag = np.sqrt(abx * abx + aby * aby + abz * abz)
    with np.errstate(divide="ignore"):   # the quotient is used only where bmag > bmax
        sf = np.where(bmag > bmax, reynolds * (bmax / bmag), reynolds).astype(F)

    cxy, cxz, cyz = wxy[C], wxz[C], wyz[C]
    cx, cy, cz = wx[C], wy[C], wz[C]
    qpresx[b][C] = (px0 - two * delt * cxy * (qvx[YP] - qvx[YM]) / dyt
                    - two * delt * cxz * (qvx[ZP] - qvx[ZM]) / dzt
                    + two * delt * sf / rmass * (cxy * abz - cxz * aby))
    qpresy[b][C] = (py0 - two * delt * cxy * (qvy[XP] - qvy[XM]) / dxt
                    - two * delt * cyz * (qvy[ZP] - qvy[ZM]) / dzt
                    + two * delt * sf / rmass * (cyz * abx - cxy * abz))
    qpresz[b][C] = (pz0 - two * delt * cxz * (qvz[XP] - qvz[XM]) / dxt
                    - two * delt * cyz * (qvz[YP] - qvz[YM]) / dyt
                    + two * delt * sf / rmass * (cxz * aby - cyz * abx))

    # off-diagonal pressures (:285-356)
    pxy = oldqpresxy[b][C] - delt * flux(wxy)
    qpresxy[b][C] = (pxy - delt * cxy * ((qvx[XP] - qvx[XM]) / dxt + (qvy[YP] - qvy[YM]) / dyt)
                     - delt * cy * (qvx[YP] - qvx[YM]) / dyt
                     - delt * cx * (qvy[XP] - qvy[XM]) / dxt
                     - delt * cyz * (qvx[ZP] - qvx[ZM]) / dzt
                     - delt * cxz * (qvy[ZP] - qvy[ZM]) / dzt
                     + delt * sf / rmass * (abz * (cy - cx) - aby * cyz + abx * cxz))
    pxz = oldqpresxz[b][C] - delt * flux(wxz)
    qpresxz[b][C] = (pxz - delt * cxz * ((qvx[XP] - qvx[XM]) / dxt + (qvz[ZP] - qvz[ZM]) / dzt)
                     - delt * cz * (qvx[ZP] - qvx[ZM]) / dzt
                     - delt * cx * (qvz[XP] - qvz[XM]) / dxt
                     - delt * cyz * (qvx[YP] - qvx[YM]) / dyt
                     - delt * cxy * (qvz[YP] - qvz[YM]) / dyt
                     + delt * sf / rmass * (aby * (cx - cz) - abx * cxy + abz * cyz))
    pyz = oldqpresyz[b][C] - delt * flux(wyz)
    qpresyz[b][C] = (pyz - delt * cyz * ((qvy[YP] - qvy[YM]) / dyt + (qvz[ZP] - qvz[ZM]) / dzt)
                     - delt * cy * (qvz[YP] - qvz[YM]) / dyt
                     - delt * cz * (qvy[ZP] - qvy[ZM]) / dzt
                     - delt * cxy * (qvz[XP] - qvz[XM]) / dxt
                     - delt * cxz * (qvy[XP] - qvy[XM]) / dxt
                     + delt * sf / rmass * (abx * (cz - cy) - abz * cxz + aby * cxy))


# ---------------------------------------------------------------------------
# code/sub_set_rho.f90:5-104  density floors and negative-pressure reset


def set_rho(qrho, qpresx, qpresy, qpresz, qpresxy, qpresxz, qpresyz, rmassq,
            hrho, hpresx, hpresy, hpresz, hpresxy, hpresxz, hpresyz, rmassh,
            orho, opresx, opresy, opresz, opresxy, opresxz, opresyz, rmasso,
            epres, nx, ny, nz, n_grids, box, o_conc):
    b = box - 1
    mins = (F(2.e-4), F(1.e-4), F(2.e-5))
    groups = ((qrho, qpresx, qpresy, qpresz, qpresxy, qpresxz, qpresyz),
              (hrho, hpresx, hpresy, hpresz, hpresxy, hpresxz, hpresyz),
              (orho, opresx, opresy, opresz, opresxy, opresxz, opresyz))
    for (r, px, py, pz, pxy, pxz, pyz), dmin in zip(groups, mins):
        R, X, Y, Z = r[b], px[b], py[b], pz[b]
        m = (R < dmin) | (X < 0) | (Y < 0) | (Z < 0)
        if not m.any():
            continue
        floor = F(1.05) * dmin
        R[m] = np.maximum(np.abs(R[m]), floor)
        ap = (np.abs(X[m]) + np.abs(Y[m]) + np.abs(Z[m])) / F(3)
        X[m] = ap
        Y[m] = ap
        Z[m] = ap
        pxy[b][m] = 0
        pxz[b][m] = 0
        pyz[b][m] = 0
    np.abs(epres[b], out=epres[b])


# ---------------------------------------------------------------------------
# code/sub_set_speed_agrd.f90:6-435  speed limits, neighbour averaging, CFL maxima

_ISO_FLAG = [F(1)]   # the reference's SAVEd iso_flag (:68,75)


def set_speed_agrd(nx, ny, nz, n_grids, box, rmassq, rmassh, rmasso,
                   isotropic, smallbit, aniso_limit, vlim, qvlim, hvlim, ovlim,
                   qclim, hclim, oclim, cslim,
                   bx, by, bz, bx0, by0, bz0,
                   qrho, qpx, qpy, qpz, qpresx, qpresy, qpresz, qpresxy, qpresxz, qpresyz,
                   hrho, hpx, hpy, hpz, hpresx, hpresy, hpresz, hpresxy, hpresxz, hpresyz,
                   orho, opx, opy, opz, opresx, opresy, opresz, opresxy, opresxz, opresyz,
                   epres, btot, pxmax, pymax, pzmax, csmax, alfmax, fastest):
    b = box - 1
    smallbit, aniso, vlim, cslim = _f(smallbit, aniso_limit, vlim, cslim)
    if isotropic:
        _ISO_FLAG[0] = F(0)
    iso = _ISO_FLAG[0]
    mx = [F(0)] * 4   # pxmax, pymax, pzmax, csmax
    species = ((qrho, qpx, qpy, qpz, qpresx, qpresy, qpresz, qpresxy, qpresxz, qpresyz, F(qvlim), F(qclim)),
               (hrho, hpx, hpy, hpz, hpresx, hpresy, hpresz, hpresxy, hpresxz, hpresyz, F(hvlim), F(hclim)),
               (orho, opx, opy, opz, opresx, opresy, opresz, opresxy, opresxz, opresyz, F(ovlim), F(oclim)))
    for (r, px, py, pz, Px, Py, Pz, Pxy, Pxz, Pyz, svl, scl) in species:
        R, X, Y, Z = r[b], px[b], py[b], pz[b]
        PX, PY, PZ = Px[b], Py[b], Pz[b]
        # snapshot with anisotropy limit (:88-121)
        s_rho, s_px, s_py, s_pz = R.copy(), X.copy(), Y.copy(), Z.copy()
        ap = (PX + PY + PZ) / F(3.0)
        one_m = F(1.) - aniso
        s_prx = aniso * PX + one_m * ap
        s_pry = aniso * PY + one_m * ap
        s_prz = aniso * PZ + one_m * ap
        s_pxy = aniso * Pxy[b]
        s_pxz = aniso * Pxz[b]
        s_pyz = aniso * Pyz[b]
        if isotropic:
            s_prx = PX.copy()
            s_pry = PX.copy()
            s_prz = PX.copy()
            s_pxy = np.zeros_like(PX)
            s_pxz = np.zeros_like(PX)
            s_pyz = np.zeros_like(PX)
        # checks on the interior (:123-184)
        arho = R[C] + smallbit
        avx = np.abs(X[C] / arho)
        avy = np.abs(Y[C] / arho)
        avz = np.abs(Z[C] / arho)
        v = np.sqrt(avx * avx + avy * avy + avz * avz)
        cs = np.sqrt(F(0.33) * (PX[C] + PY[C] + PZ[C]) / arho)
        flag = (v > (vlim * svl)) | (cs > vlim * svl * scl)
        if flag.any():
            def avg(s):
                return (s[XP] + s[XM] + s[YP] + s[YM] + s[ZP] + s[ZM]) / F(6.)

            def avg_iso(s):
                return iso * (s[XP] + s[XM] + s[YP] + s[YM] + s[ZP] + s[ZM]) / F(6.)
            for dst, snap, fn in ((R, s_rho, avg), (X, s_px, avg), (Y, s_py, avg), (Z, s_pz, avg),
                                  (PX, s_prx, avg), (PY, s_pry, avg), (PZ, s_prz, avg),
                                  (Pxy[b], s_pxy, avg_iso), (Pxz[b], s_pxz, avg_iso),
                                  (Pyz[b], s_pyz, avg_iso)):
                interior = dst[C]
                interior[flag] = fn(snap)[flag]
        mx[0] = max(mx[0], avx.max())
        mx[1] = max(mx[1], avy.max())
        mx[2] = max(mx[2], avz.max())
        mx[3] = max(mx[3], cs.max())
    # total field and electron sound-speed cap (:385-429)
    bsx = bx0[b] + bx[b]
    bsy = by0[b] + by[b]
    bsz = bz0[b] + bz[b]
    btot[...] = np.maximum(np.sqrt(bsx * bsx + bsy * bsy + bsz * bsz), smallbit)
    arho = qrho[b][C] + hrho[b][C] + orho[b][C] + smallbit
    E = epres[b]
    cs = np.sqrt(E[C] / arho)
    cap = cs > cslim
    if cap.any():
        ec = E[C]
        ec[cap] = ec[cap] * cslim / cs[cap]
    amax = (btot[C] / np.sqrt(arho)).max()
    pxmax[0], pymax[0], pzmax[0], csmax[0] = mx
    alfmax[0] = max(F(0), amax)
    fastest[0] = max(F(0), np.sqrt(mx[0] * mx[0] + mx[1] * mx[1] + mx[2] * mx[2]
                                   + mx[3] * mx[3] + alfmax[0] * alfmax[0]))


# ---------------------------------------------------------------------------
# Lapidus smoothing: code/sub_lap_plasma.f90:4-147, sub_lap_elec.f90:4-51,
# sub_lap_bfld.f90:4-67


def _lap(q, w, vx, vy, vz, chi, delt):
    uxp = np.abs(vx[XP] - vx[C])
    uxm = np.abs(vx[C] - vx[XM])
    uyp = np.abs(vy[YP] - vy[C])
    uym = np.abs(vy[C] - vy[YM])
    uzp = np.abs(vz[ZP] - vz[C])
    uzm = np.abs(vz[C] - vz[ZM])
    c = q[C]
    w[C] = c + chi * (delt * (uxp * (q[XP] - c) - uxm * (c - q[XM]))
                      + delt * (uyp * (q[YP] - c) - uym * (c - q[YM]))
                      + delt * (uzp * (q[ZP] - c) - uzm * (c - q[ZM])))


def lap_plasma(rho, px, py, pz, presx, presy, presz, presxy, presxz, presyz,
               wrkrho, wrkpx, wrkpy, wrkpz, wrkpresx, wrkpresy, wrkpresz,
               wrkpresxy, wrkpresxz, wrkpresyz, vx, vy, vz,
               nx, ny, nz, n_grids, box, chirho, chipxyz, chierg, delt, isotropic):
    b = box - 1
    chirho, chipxyz, delt = _f(chirho, chipxyz, delt)
    _lap(rho[b], wrkrho[b], vx, vy, vz, chirho, delt)
    for q, w in ((px, wrkpx), (py, wrkpy), (pz, wrkpz), (presx, wrkpresx),
                 (presy, wrkpresy), (presz, wrkpresz), (presxy, wrkpresxy),
                 (presxz, wrkpresxz), (presyz, wrkpresyz)):
        _lap(q[b], w[b], vx, vy, vz, chipxyz, delt)


def lap_elec(ppres, wrkppres, vx, vy, vz, nx, ny, nz, n_grids, box,
             chirho, chipxyz, chierg, delt):
    b = box - 1
    _lap(ppres[b], wrkppres[b], vx, vy, vz, F(chipxyz), F(delt))


def lap_bfld(bx, by, bz, wrkbx, wrkby, wrkbz, vx, vy, vz, nx, ny, nz, n_grids, box,
             chirho, chipxyz, chierg, delt):
    b = box - 1
    for q, w in ((bx, wrkbx), (by, wrkby), (bz, wrkbz)):
        _lap(q[b], w[b], vx, vy, vz, F(chierg), F(delt))


# code/sub_fnd_vtot.f90:4-41
def fnd_vtot(qpx, qpy, qpz, qrho, hpx, hpy, hpz, hrho, opx, opy, opz, orho,
             vx, vy, vz, nx, ny, nz, n_grids, box, rmassq, rmassh, rmasso):
    b = box - 1
    tden = qrho[b] + hrho[b] + orho[b]
    vx[...] = (qpx[b] + hpx[b] + opx[b]) / tden
    vy[...] = (qpy[b] + hpy[b] + opy[b]) / tden
    vz[...] = (qpz[b] + hpz[b] + opz[b]) / tden


# ---------------------------------------------------------------------------
# code/sub_fcsmooth.f90:7-242  flux-corrected transport, one field


def _limit(s, d1m, chiabs, d1p):
    return s * np.maximum(F(0), np.minimum(np.minimum(s * d1m, chiabs), s * d1p))


def fcsmooth(px, oldpx, wrkpx, nx, ny, nz, n_grids, box, chipx, tempx, tempy, tempz):
    b = box - 1
    chipx = F(chipx)
    P, O, W = px[b], oldpx[b], wrkpx[b]
    tempx[...] = W
    # diffusion (:23-33)
    P[C] = W[C] + chipx * (O[XP] + O[XM] + O[YP] + O[YM] + O[ZP] + O[ZM] - F(6.) * O[C])
    # boundary points (:37-65)
    P[:, :, 0] = W[:, :, 0]
    P[:, :, -1] = W[:, :, -1]
    P[0] = W[0]
    P[-1] = W[-1]
    P[:, 0, :] = W[:, 0, :]
    P[:, -1, :] = W[:, -1, :]
    # x antidiffusion (:69-116)
    dx = np.empty_like(W)
    d1x = np.empty_like(W)
    dx[:, :, :-1] = W[:, :, 1:] - W[:, :, :-1]
    d1x[:, :, :-1] = P[:, :, 1:] - P[:, :, :-1]
    dx[:, :, -1] = dx[:, :, -2]
    d1x[:, :, -1] = d1x[:, :, -2]
    fn = np.empty_like(W)
    s = np.copysign(F(1), dx[:, :, 1:-1])
    fn[:, :, 1:-1] = _limit(s, d1x[:, :, :-2], chipx * np.abs(dx[:, :, 1:-1]), d1x[:, :, 2:])
    fn[:, :, 0] = fn[:, :, 1]
    fn[:, :, -1] = fn[:, :, -2]
    tempx[:, :, 1:-1] = P[:, :, 1:-1] - fn[:, :, 1:-1] + fn[:, :, :-2]
    # y antidiffusion (:120-167)
    dy = np.empty_like(W)
    d1y = np.empty_like(W)
    dy[:, :-1, :] = W[:, 1:, :] - W[:, :-1, :]
    d1y[:, :-1, :] = P[:, 1:, :] - P[:, :-1, :]
    dy[:, -1, :] = dy[:, -2, :]
    d1y[:, -1, :] = d1y[:, -2, :]
    s = np.copysign(F(1), dy[:, 1:-1, :])
    fn[:, 1:-1, :] = _limit(s, d1y[:, :-2, :], chipx * np.abs(dy[:, 1:-1, :]), d1y[:, 2:, :])
    fn[:, 0, :] = fn[:, 1, :]
    fn[:, -1, :] = fn[:, -2, :]
    tempx[:, 1:-1, :] = tempx[:, 1:-1, :] - fn[:, 1:-1, :] + fn[:, :-2, :]
    # z antidiffusion (:171-239)
    dz = np.empty_like(W)
    d1z = np.empty_like(W)
    dz[:-1] = W[1:] - W[:-1]
    d1z[:-1] = P[1:] - P[:-1]
    dz[-1] = dz[-2]
    d1z[-1] = d1z[-2]
    s = np.copysign(F(1), dz[1:-1])
    fn[1:-1] = _limit(s, d1z[:-2], chipx * np.abs(dz[1:-1]), d1z[2:])
    fn[0] = fn[1]
    fn[-1] = fn[-2]
    P[1:-1] = tempx[1:-1] - fn[1:-1] + fn[:-2]


def flux_correct(qrho, qpresx, qpresy, qpresz, qpresxy, qpresxz, qpresyz, qpx, qpy, qpz,
                 wrkqrho, wrkqpresx, wrkqpresy, wrkqpresz, wrkqpresxy, wrkqpresxz, wrkqpresyz,
                 wrkqpx, wrkqpy, wrkqpz,
                 oldqrho, oldqpresx, oldqpresy, oldqpresz, oldqpresxy, oldqpresxz, oldqpresyz,
                 oldqpx, oldqpy, oldqpz,
                 hrho, hpresx, hpresy, hpresz, hpresxy, hpresxz, hpresyz, hpx, hpy, hpz,
                 wrkhrho, wrkhpresx, wrkhpresy, wrkhpresz, wrkhpresxy, wrkhpresxz, wrkhpresyz,
                 wrkhpx, wrkhpy, wrkhpz,
                 oldhrho, oldhpresx, oldhpresy, oldhpresz, oldhpresxy, oldhpresxz, oldhpresyz,
                 oldhpx, oldhpy, oldhpz,
                 orho, opresx, opresy, opresz, opresxy, opresxz, opresyz, opx, opy, opz,
                 wrkorho, wrkopresx, wrkopresy, wrkopresz, wrkopresxy, wrkopresxz, wrkopresyz,
                 wrkopx, wrkopy, wrkopz,
                 oldorho, oldopresx, oldopresy, oldopresz, oldopresxy, oldopresxz, oldopresyz,
                 oldopx, oldopy, oldopz,
                 epres, wrkepres, oldepres, bx, by, bz, wrkbx, wrkby, wrkbz, oldbx, oldby, oldbz,
                 vvx, vvy, vvz, nx, ny, nz, n_grids, box, difrho, diferg, xspac, isotropic):
    """code/sub_flux_correct.f90:4-192: fcsmooth on all 34 fields."""
    a = locals()
    chi = F(difrho)
    for sp in "qho":
        for nm in ("rho", "presx", "presy", "presz", "presxy", "presxz", "presyz", "px", "py", "pz"):
            fcsmooth(a[sp + nm], a["old" + sp + nm], a["wrk" + sp + nm],
                     nx, ny, nz, n_grids, box, chi, vvx, vvy, vvz)
    fcsmooth(epres, oldepres, wrkepres, nx, ny, nz, n_grids, box, chi, vvx, vvy, vvz)
    chi = F(diferg)
    for c in "xyz":
        fcsmooth(a["b" + c], a["oldb" + c], a["wrkb" + c], nx, ny, nz, n_grids, box, chi,
                 vvx, vvy, vvz)


# ---------------------------------------------------------------------------
# code/sub_bndry_outer.f90:5-731  outer-box faces, edges and corners

# kinds: which sign rule applies on which face
_RHO, _PRES, _PX, _PY, _PZ, _OFF, _EP, _BX, _BY, _BZ = range(10)


def bndry_outer(qrho, qpresx, qpresy, qpresz, qpx, qpy, qpz,
                hrho, hpresx, hpresy, hpresz, hpx, hpy, hpz,
                orho, opresx, opresy, opresz, opx, opy, opz,
                epres, qpresxy, qpresxz, qpresyz, hpresxy, hpresxz, hpresyz,
                opresxy, opresxz, opresyz, rmassq, rmassh, rmasso,
                bx, by, bz, bx0, by0, bz0, nx, ny, nz, n_grids,
                srho, rho_frac, o_conc, spress, spx, spy, spz,
                sbx_wind, sby_wind, sbz_wind, ti_te, isotropic):
    b = n_grids - 1
    rmassq, rmassh, rmasso = _f(rmassq, rmassh, rmasso)
    srho, rho_frac, o_conc, spress, spx, spy, spz = _f(srho, rho_frac, o_conc, spress, spx, spy, spz)
    sbx, sby, sbz, ti_te = _f(sbx_wind, sby_wind, sbz_wind, ti_te)
    ofrac = rho_frac * o_conc
    frac_h = rmassh / rmassq
    frac_o = rmasso / rmassq
    half = F(0.5)
    pq = half * spress
    # (field, kind, inflow value or None) -- inflow values (:50-103)
    fields = [
        (qrho, _RHO, srho), (qpx, _PX, spx), (qpy, _PY, spy), (qpz, _PZ, spz),
        (qpresx, _PRES, pq), (qpresy, _PRES, pq), (qpresz, _PRES, pq),
        (hrho, _RHO, srho * rho_frac * frac_h), (hpx, _PX, spx * rho_frac * frac_h),
        (hpy, _PY, spy * rho_frac * frac_h), (hpz, _PZ, spz * rho_frac * frac_h),
        (hpresx, _PRES, half * spress * rho_frac), (hpresy, _PRES, half * spress * rho_frac),
        (hpresz, _PRES, half * spress * rho_frac),
        (orho, _RHO, srho * ofrac * frac_o), (opx, _PX, spx * ofrac * frac_o),
        (opy, _PY, spy * ofrac * frac_o), (opz, _PZ, spz * ofrac * frac_o),
        (opresx, _PRES, half * spress * ofrac), (opresy, _PRES, half * spress * ofrac),
        (opresz, _PRES, half * spress * ofrac),
        (epres, _EP, pq / ti_te),
        (qpresxy, _OFF, F(0)), (qpresxz, _OFF, F(0)), (qpresyz, _OFF, F(0)),
        (hpresxy, _OFF, F(0)), (hpresxz, _OFF, F(0)), (hpresyz, _OFF, F(0)),
        (opresxy, _OFF, F(0)), (opresxz, _OFF, F(0)), (opresyz, _OFF, F(0)),
        (bx, _BX, None), (by, _BY, None), (bz, _BZ, None),
    ]
    two = F(2.)
    for arr, kind, val in fields:
        f = arr[b]
        # inflow face i=1
        if kind == _BX:
            f[:, :, 0] = sbx - bx0[b][:, :, 0]
        elif kind == _BY:
            f[:, :, 0] = sby - by0[b][:, :, 0]
        elif kind == _BZ:
            f[:, :, 0] = sbz - bz0[b][:, :, 0]
        else:
            f[:, :, 0] = val
        # back wall i=nx (:107-154)
        f[:, :, -1] = np.abs(f[:, :, -2]) if kind == _PX else f[:, :, -2]
        # y faces, i=2..nx (:158-252)
        if kind == _PY:
            f[:, 0, 1:] = -np.abs(f[:, 1, 1:])
            f[:, -1, 1:] = np.abs(f[:, -2, 1:])
        else:
            f[:, 0, 1:] = f[:, 1, 1:]
            f[:, -1, 1:] = f[:, -2, 1:]
        # z top, i=2..nx (:256-307)
        f[-1, :, 1:] = np.abs(f[-2, :, 1:]) if kind == _PZ else f[-2, :, 1:]
        # z bottom (:311-359)
        f[0, :, 1:] = -np.abs(f[1, :, 1:]) if kind == _PZ else f[1, :, 1:]
        # corner lines of the back wall, k=2..nz-1 (:363-466)
        if kind == _BX:
            f[1:-1, 0, -1] = f[1:-1, 0, -2]
            f[1:-1, -1, -1] = f[1:-1, -1, -2]
        else:
            f[1:-1, 0, -1] = (f[1:-1, 0, -2] + f[1:-1, 1, -1]) / two
            f[1:-1, -1, -1] = (f[1:-1, -1, -2] + f[1:-1, -2, -1]) / two
        # top and bottom lines of the back wall, j=2..ny-1 (:470-566)
        f[0, 1:-1, -1] = f[1, 1:-1, -1]
        if kind != _EP:
            f[-1, 1:-1, -1] = f[-1, 1:-1, -2]
        # x-directed edge lines, i=2..nx (:570-728)
        f[0, 0, 1:] = f[1, 0, 1:]
        f[0, -1, 1:] = f[1, -1, 1:]
        f[-1, 0, 1:] = f[-2, 0, 1:]
        f[-1, -1, 1:] = f[-2, -1, 1:]


# ---------------------------------------------------------------------------
# code/sub_bndry_inner.f90:5-268  ionosphere / atmosphere / interior lists


def bndry_inner(qrho, qpresx, qpresy, qpresz, qpx, qpy, qpz, rmassq,
                hrho, hpresx, hpresy, hpresz, hpx, hpy, hpz, rmassh,
                orho, opresx, opresy, opresz, opx, opy, opz, rmasso,
                epres, bx, by, bz, qpresxy, qpresxz, qpresyz,
                hpresxy, hpresxz, hpresyz, opresxy, opresxz, opresyz,
                nx, ny, nz, n_grids, parm_srf, parm_mid, parm_zero,
                ijsrf, numsrf, ijmid, nummid, ijzero, numzero,
                mbndry, msrf, mmid, mzero, erho, epress, alpha_e, ti_te, o_conc,
                re_equiv, r_inner, sbx_wind, spress,
                grd_xmin, grd_xmax, grd_ymin, grd_ymax, grd_zmin, grd_zmax,
                v_rot=0.0, r_rot=0.0):
    """``v_rot``/``r_rot`` come from the reference's common /rotation/."""
    b = mbndry - 1
    re_equiv, v_rot = _f(re_equiv, v_rot)
    dx = (F(grd_xmax[b]) - F(grd_xmin[b])) / F(nx - 1.)
    dy = (F(grd_ymax[b]) - F(grd_ymin[b])) / F(ny - 1.)
    one = F(1.)
    rhos = (qrho[b], hrho[b], orho[b])
    pres = ((qpresx[b], qpresy[b], qpresz[b]), (hpresx[b], hpresy[b], hpresz[b]),
            (opresx[b], opresy[b], opresz[b]))
    moms = ((qpx[b], qpy[b], qpz[b]), (hpx[b], hpy[b], hpz[b]), (opx[b], opy[b], opz[b]))
    offs = (qpresxy[b], qpresxz[b], qpresyz[b], hpresxy[b], hpresxz[b], hpresyz[b],
            opresxy[b], opresxz[b], opresyz[b])

    def coords(lst, n):
        ii = lst[:n, 0, b] - 1
        jj = lst[:n, 1, b] - 1
        kk = lst[:n, 2, b] - 1
        ay = F(grd_ymin[b]) + dy * jj.astype(F)
        ax = F(grd_xmin[b]) + dx * ii.astype(F)
        rx = ax * re_equiv
        ry = ay * re_equiv
        rvy = rx * v_rot * one
        rvx = -ry * v_rot * one
        return (kk, jj, ii), rvx, rvy

    # surface (:78-140): densities floored at the stored values
    n = int(numsrf[b])
    if n:
        idx, rvx, rvy = coords(ijsrf, n)
        p = parm_srf[:n, :, b]
        dens = [np.maximum(p[:, s], rhos[s][idx]) for s in range(3)]
        for s in range(3):
            rhos[s][idx] = dens[s]
        for s in range(3):
            for c in range(3):
                pres[s][c][idx] = p[:, 3 + s]
        epres[b][idx] = p[:, 6]
        for s in (1, 2, 0):
            moms[s][0][idx] = dens[s] * rvx
            moms[s][1][idx] = dens[s] * rvy
            moms[s][2][idx] = dens[s] * F(0)
        for o in offs:
            o[idx] = 0
    # atmosphere (:143-201) and interior (:204-265)
    for lst, num, parm, zero_b in ((ijmid, nummid, parm_mid, False),
                                   (ijzero, numzero, parm_zero, True)):
        n = int(num[b])
        if not n:
            continue
        idx, rvx, rvy = coords(lst, n)
        p = parm[:n, :, b]
        for s in range(3):
            rhos[s][idx] = p[:, s]
        for s in range(3):
            for c in range(3):
                pres[s][c][idx] = p[:, 3 + s]
        epres[b][idx] = p[:, 6]
        for s in (2, 1, 0):
            r = rhos[s][idx]
            moms[s][0][idx] = r * rvx
            moms[s][1][idx] = r * rvy
            moms[s][2][idx] = r * F(0)
        for o in offs:
            o[idx] = 0
        if zero_b:
            bx[b][idx] = 0
            by[b][idx] = 0
            bz[b][idx] = 0


# ---------------------------------------------------------------------------
# Nested-grid coupling


def _axis_weights(gmin_s, gmax_s, gmin_b, gmax_b, n):
    """Coarse index and weights of each fine index along one axis
    (code/sub_flanks.f90:24-30, 50-55)."""
    dl = (F(gmax_b) - F(gmin_b)) / F(n - 1.)
    sl = (F(gmax_s) - F(gmin_s)) / F(n - 1.)
    a = F(gmin_s) + sl * np.arange(n, dtype=F)
    ai = F(1.) + (a - F(gmin_b)) / dl
    ib = ai.astype(np.int64)              # truncation toward zero
    d = ai - ib.astype(F)
    return ib - 1, d, F(1.) - d           # 0-based lower index


def flanks(wrkrho, rho, oldrho, nx, ny, nz, n_grids, box, t_new, t_old, t, work,
           grd_xmin, grd_xmax, grd_ymin, grd_ymax, grd_zmin, grd_zmax):
    """code/sub_flanks.f90:5-303: coarse->fine boundary, 2 layers per face,
    trilinear in space and linear in time; panels in reference order."""
    mb, ms = box, box - 1          # 0-based coarse / fine
    t = float(t.v if hasattr(t, "v") else t)
    t_new = np.asarray(t_new, dtype=np.float64)
    t_old = np.asarray(t_old, dtype=np.float64)
    t1 = F((t_new[mb] - t) / (t_new[mb] - t_old[mb]))
    t2 = F(1.) - t1
    work[...] = oldrho[mb] * t1 + rho[mb] * t2
    ib, dxv, ddxv = _axis_weights(grd_xmin[ms], grd_xmax[ms], grd_xmin[mb], grd_xmax[mb], nx)
    jb, dyv, ddyv = _axis_weights(grd_ymin[ms], grd_ymax[ms], grd_ymin[mb], grd_ymax[mb], ny)
    kb, dzv, ddzv = _axis_weights(grd_zmin[ms], grd_zmax[ms], grd_zmin[mb], grd_zmax[mb], nz)
    W = wrkrho[ms]

    def interp(ks, js, is_):
        K, J, I = np.ix_(ks, js, is_)
        k0, j0, i0 = kb[K], jb[J], ib[I]
        dz, ddz = dzv[K], ddzv[K]
        dy, ddy = dyv[J], ddyv[J]
        dx, ddx = dxv[I], ddxv[I]
        w = work
        return (w[k0, j0, i0] * ddx * ddy * ddz
                + w[k0 + 1, j0, i0] * ddx * ddy * dz
                + w[k0, j0 + 1, i0] * ddx * dy * ddz
                + w[k0 + 1, j0 + 1, i0] * ddx * dy * dz
                + w[k0, j0, i0 + 1] * dx * ddy * ddz
                + w[k0 + 1, j0, i0 + 1] * dx * ddy * dz
                + w[k0, j0 + 1, i0 + 1] * dx * dy * ddz
                + w[k0 + 1, j0 + 1, i0 + 1] * dx * dy * dz)

    allk, allj, alli = np.arange(nz), np.arange(ny), np.arange(nx)

    # Layer l=0 is the outer layer (frac 0: replaced), l=1 the inner one
    # (frac 0.5: blended).  Panels are applied in the reference's order --
    # bottom, top, front, back, left, right -- so edge cells see the same
    # sequence of blends.
    def blend(sel_k, sel_j, sel_i, frac):
        frac = F(frac)
        frac1 = F(1.) - frac
        K, J, I = np.ix_(sel_k, sel_j, sel_i)
        W[K, J, I] = frac * W[K, J, I] + frac1 * interp(sel_k, sel_j, sel_i)

    for l in (1, 0):
        blend([l], allj, alli, l / 2.)
    for l in (1, 0):
        blend([nz - 1 - l], allj, alli, l / 2.)
    for l in (1, 0):
        blend(allk, [l], alli, l / 2.)
    for l in (1, 0):
        blend(allk, [ny - 1 - l], alli, l / 2.)
    for l in (1, 0):
        blend(allk, allj, [l], l / 2.)
    for l in (1, 0):
        blend(allk, allj, [nx - 1 - l], l / 2.)


def bndry_flanks(*args):
    """code/sub_bndry_flanks.f90:5-212: ``flanks`` on all 34 fields.

    Positional layout as the reference: 34 targets (wrk), 34 current (q),
    34 old, then work, nx, ny, nz, n_grids, box, t_old, t_new, t, grid."""
    tgt, cur, old = args[0:34], args[34:68], args[68:102]
    work, nx, ny, nz, n_grids, box, t_old, t_new, t = args[102:111]
    grid = args[111:117]
    # the reference calls flanks field by field in this same order
    for w, q, o in zip(tgt, cur, old):
        flanks(w, q, o, nx, ny, nz, n_grids, box, t_new, t_old, t, work, *grid)


def corer(rho, nx, ny, nz, n_grids, box, grd_xmin, grd_xmax, grd_ymin, grd_ymax,
          grd_zmin, grd_zmax):
    """code/sub_corer.f90:5-55: fine->coarse core restriction."""
    mb, ms = box, box - 1
    sx = (F(grd_xmax[ms]) - F(grd_xmin[ms])) / F(nx - 1.)
    dx = (F(grd_xmax[mb]) - F(grd_xmin[mb])) / F(nx - 1.)
    dy = (F(grd_ymax[mb]) - F(grd_ymin[mb])) / F(ny - 1.)
    dz = (F(grd_zmax[mb]) - F(grd_zmin[mb])) / F(nz - 1.)
    nskip = int(dx / sx)
    nbx1 = int(F(1.) + (F(grd_xmin[ms]) - F(grd_xmin[mb])) / dx + F(.5))
    nby1 = int(F(1.) + (F(grd_ymin[ms]) - F(grd_ymin[mb])) / dy + F(.5))
    nbz1 = int(F(1.) + (F(grd_zmin[ms]) - F(grd_zmin[mb])) / dz + F(.5))
    ks = np.arange(1 + nskip, nz - nskip + 1, nskip)   # 1-based
    js = np.arange(1 + nskip, ny - nskip + 1, nskip)
    is_ = np.arange(1 + nskip, nx - nskip + 1, nskip)
    kb = nbz1 + (ks - 1) // nskip
    jb = nby1 + (js - 1) // nskip
    ib = nbx1 + (is_ - 1) // nskip
    f = rho[ms]
    K, J, I = np.ix_(ks - 1, js - 1, is_ - 1)
    val = (F(6.) * f[K, J, I] + f[K, J, I + 1] + f[K, J, I - 1] + f[K, J + 1, I]
           + f[K, J - 1, I] + f[K + 1, J, I] + f[K - 1, J, I]) / F(12.)
    rho[mb][np.ix_(kb - 1, jb - 1, ib - 1)] = val


_CORER_ORDER = None


def bndry_corer(qrho, qpresx, qpresy, qpresz, qpx, qpy, qpz,
                hrho, hpresx, hpresy, hpresz, hpx, hpy, hpz,
                orho, opresx, opresy, opresz, opx, opy, opz,
                epres, bx, by, bz, qpresxy, qpresxz, qpresyz,
                hpresxy, hpresxz, hpresyz, opresxy, opresxz, opresyz,
                nx, ny, nz, n_grids, box, *grid):
    """code/sub_bndry_corer.f90:5-114: ``corer`` on all 34 fields."""
    for f in (qrho, qpresx, qpresy, qpresz, qpresxy, qpresxz, qpresyz, qpx, qpy, qpz,
              hrho, hpresx, hpresy, hpresz, hpresxy, hpresxz, hpresyz, hpx, hpy, hpz,
              orho, opresx, opresy, opresz, opresxy, opresxz, opresyz, opx, opy, opz,
              epres, bx, by, bz):
        corer(f, nx, ny, nz, n_grids, box, *grid)


def flanks_synced(rho, nx, ny, nz, n_grids, box, grd_xmin, grd_xmax, grd_ymin,
                  grd_ymax, grd_zmin, grd_zmax):
    """code/sub_flanks_synced.f90:5-141: single-layer coarse->fine faces,
    bilinear; top/bottom, front/back, left/right in order."""
    mb, ms = box, box - 1
    delx = (F(grd_xmax[mb]) - F(grd_xmin[mb])) / F(nx - 1.)
    dely = (F(grd_ymax[mb]) - F(grd_ymin[mb])) / F(ny - 1.)
    delz = (F(grd_zmax[mb]) - F(grd_zmin[mb])) / F(nz - 1.)
    kb_top = int(F(1.) + (F(grd_zmax[ms]) - F(grd_zmin[mb])) / delz) - 1
    kb_bot = int(F(1.) + (F(grd_zmin[ms]) - F(grd_zmin[mb])) / delz) - 1
    jb_top = int(F(1.) + (F(grd_ymax[ms]) - F(grd_ymin[mb])) / dely) - 1
    jb_bot = int(F(1.) + (F(grd_ymin[ms]) - F(grd_ymin[mb])) / dely) - 1
    ib_top = int(F(1.) + (F(grd_xmax[ms]) - F(grd_xmin[mb])) / delx) - 1
    ib_bot = int(F(1.) + (F(grd_xmin[ms]) - F(grd_xmin[mb])) / delx) - 1
    ib, dx, ddx = _axis_weights(grd_xmin[ms], grd_xmax[ms], grd_xmin[mb], grd_xmax[mb], nx)
    jb, dy, ddy = _axis_weights(grd_ymin[ms], grd_ymax[ms], grd_ymin[mb], grd_ymax[mb], ny)
    kb, dz, ddz = _axis_weights(grd_zmin[ms], grd_zmax[ms], grd_zmin[mb], grd_zmax[mb], nz)
    S, B = rho[ms], rho[mb]
    # top and bottom panels (:43-72): arrays over (js, is)
    J, I = np.ix_(jb, ib)
    DY, DDY = dy[:, None], ddy[:, None]
    DX, DDX = dx[None, :], ddx[None, :]
    for ks, kbb in ((-1, kb_top), (0, kb_bot)):
        S[ks] = (B[kbb, J, I] * DDX * DDY + B[kbb, J + 1, I] * DDX * DY
                 + B[kbb, J, I + 1] * DX * DDY + B[kbb, J + 1, I + 1] * DX * DY)
    # front and back panels (:76-105): arrays over (ks, is)
    K, I = np.ix_(kb, ib)
    DZ, DDZ = dz[:, None], ddz[:, None]
    for js, jbb in ((-1, jb_top), (0, jb_bot)):
        S[:, js, :] = (B[K, jbb, I] * DDX * DDZ + B[K + 1, jbb, I] * DDX * DZ
                       + B[K, jbb, I + 1] * DX * DDZ + B[K + 1, jbb, I + 1] * DX * DZ)
    # left and right panels (:109-138): arrays over (ks, js)
    K, J = np.ix_(kb, jb)
    DY2, DDY2 = dy[None, :], ddy[None, :]
    for is_, ibb in ((-1, ib_top), (0, ib_bot)):
        S[:, :, is_] = (B[K, J, ibb] * DDY2 * DDZ + B[K, J + 1, ibb] * DY2 * DDZ
                        + B[K + 1, J, ibb] * DDY2 * DZ + B[K + 1, J + 1, ibb] * DY2 * DZ)


# ---------------------------------------------------------------------------
# divB clean-up: code/sub_banded_tpose.f90:6-33, sub_sparse_solver.f90:19-159,
# sub_divB_correct.f90:25-175


def seq_sum(x):
    """Sequential single-precision sum, as the reference's DO loops add."""
    return np.add.accumulate(x.astype(F), dtype=F)[-1] if x.size else F(0)


def banded_tpose(nbands, npts, divB_band_pos, banded_matrix, x_vec, tp_result):
    """y_i = sum_b M(b, i+pos_b) x(i+pos_b), accumulated in band order."""
    M = banded_matrix.reshape(npts, nbands)      # Fortran M(nbands, npts)
    y = np.zeros(npts, dtype=F)
    idx = np.arange(npts)
    for band in range(nbands):
        ii = idx + int(divB_band_pos[band])
        ok = (ii >= 0) & (ii < npts)
        t = np.zeros(npts, dtype=F)
        t[ok] = M[ii[ok], band] * x_vec[ii[ok]]
        y = y + t
    tp_result[...] = y


_REDOS = [0]   # the reference's SAVEd `redos` (:45)


def sparse_solver(nbands, npts, divB_band_pos, divB_errlim, recurse, redo_lim,
                  divB, soln_vec, residual2, banded_matrix):
    """Numerical-Recipes CGNR with restart recursion, all single precision."""
    errlim2 = F(npts) * (F(divB_errlim) * F(divB_errlim))
    _REDOS[0] = int(redo_lim) if recurse else 0
    _REDOS[0] += 1
    xi = np.empty(npts, F)
    xj = np.empty(npts, F)
    g = np.empty(npts, F)
    banded_tpose(nbands, npts, divB_band_pos, banded_matrix, soln_vec, xi)
    sum_b = seq_sum(divB * divB)
    xi[...] = xi - divB
    resd = seq_sum(xi * xi)
    banded_tpose(nbands, npts, divB_band_pos, banded_matrix, xi, g)
    g[...] = -g
    h = g.copy()
    for _ in range(10 * npts):
        banded_tpose(nbands, npts, divB_band_pos, banded_matrix, h, xi)
        anum = seq_sum(g * h)
        aden = seq_sum(xi * xi)
        anum = anum / aden
        xi[...] = soln_vec
        soln_vec[...] = soln_vec + anum * h
        banded_tpose(nbands, npts, divB_band_pos, banded_matrix, soln_vec, xj)
        xj[...] = xj - divB
        r2 = seq_sum(xj * xj)
        residual2[0] = r2
        if r2 == resd or r2 <= sum_b * errlim2:
            return
        if r2 > resd:
            soln_vec[...] = xi
            if _REDOS[0] >= 3:
                return
            sparse_solver(nbands, npts, divB_band_pos, divB_errlim, True, _REDOS[0],
                          divB, soln_vec, residual2, banded_matrix)
        resd = residual2[0]
        banded_tpose(nbands, npts, divB_band_pos, banded_matrix, xj, xi)
        gg = seq_sum(g * g)
        dgg = seq_sum((xi + g) * xi)
        if gg == 0:
            return
        gam = dgg / gg
        g[...] = -xi
        h[...] = g + gam * h


def divb_matrix(nx, ny, nz):
    """The 7-band matrix of code/sub_divB_correct.f90:85-123, as a
    C-order (npts, 7) array (= Fortran banded_matrix(7, npts))."""
    npts = nx * ny * nz
    M = np.zeros((npts, 7), dtype=F)
    M[:, 3] = -6.
    n = np.arange(npts)
    i = n % nx
    jp = n % (nx * ny)
    M[i >= 1, 2] = 1.                 # upper1: i = 2..nx
    M[i <= nx - 2, 4] = 1.            # lower1: i = 1..nx-1
    M[jp >= nx, 1] = 1.               # upper2: in-plane index nx+1..nx*ny
    M[jp < nx * ny - nx, 5] = 1.      # lower2
    M[n >= nx * ny, 0] = 1.           # upper3
    M[n < npts - nx * ny, 6] = 1.     # lower3
    return M


def divB_correct(nx, ny, nz, n_grids, box, nbands, divB_band_pos, xspac, divB_errlim,
                 bx, by, bz, soln_init=None):
    """The reference leaves ``soln_vec`` (an automatic array) uninitialised
    before the solve; the oracle starts from ``soln_init`` (zeros by
    default), the only defined choice."""
    b = box - 1
    rx = F(xspac[b])
    npts = nx * ny * nz
    M = divb_matrix(nx, ny, nz)
    divB = np.zeros((nz, ny, nx), F)
    X, Y, Z = bx[b], by[b], bz[b]
    divB[C] = ((X[XP] - X[XM]) + (Y[YP] - Y[YM]) + (Z[ZP] - Z[ZM])) / (F(2.) * rx)
    soln = np.zeros(npts, F) if soln_init is None else soln_init.astype(F).copy()
    res = np.zeros(1, F)
    sparse_solver(nbands, npts, divB_band_pos, divB_errlim, False, 0, divB.reshape(-1),
                  soln, res, M)
    poten = (rx * rx) * soln.reshape(nz, ny, nx)
    X[C] = X[C] - (poten[XP] - poten[XM]) / (F(2.) * rx)
    Y[C] = Y[C] - (poten[YP] - poten[YM]) / (F(2.) * rx)
    Z[C] = Z[C] - (poten[ZP] - poten[ZM]) / (F(2.) * rx)
    return soln, res[0]


# ---------------------------------------------------------------------------
# code/cft_trilin_interp.f90:10-66


def trilin_interp(sxyz, gridpts, qty, scdata):
    """``gridpts`` is Fortran gridpts(3,2) -> numpy (2,3); ``qty(2,2,2)``
    -> numpy (2,2,2) indexed [z][y][x]."""
    s = np.asarray(sxyz, F)
    x0, y0, z0 = gridpts[0]
    x1, y1, z1 = gridpts[1]
    xd = (s[0] - x0) / (x1 - x0)
    yd = (s[1] - y0) / (y1 - y0)
    zd = (s[2] - z0) / (z1 - z0)
    q = qty
    one = F(1)
    e00 = q[0, 0, 0] * (one - xd) + q[0, 0, 1] * xd
    e01 = q[1, 0, 0] * (one - xd) + q[1, 0, 1] * xd
    e10 = q[0, 1, 0] * (one - xd) + q[0, 1, 1] * xd
    e11 = q[1, 1, 0] * (one - xd) + q[1, 1, 1] * xd
    f0 = e00 * (one - yd) + e10 * yd
    f1 = e01 * (one - yd) + e11 * yd
    scdata[0] = f0 * (one - zd) + f1 * zd


# ---------------------------------------------------------------------------
# code/sub_fnd_pres.f90:5-67 (pxy, pxz, pyz are read by the reference but do
# not enter the result)


def fnd_pres(px, py, pz, pxy, pxz, pyz, p_para, p_perp, p_cross, vvx, vvy, vvz, bxt, byt, bzt,
             nx, ny, nz, n_grids, box):
    b = box - 1
    x, y, z = px[b], py[b], pz[b]
    presmag = np.sqrt(x * x + y * y + z * z) + F(1.e-11)
    bmag = np.sqrt(bxt * bxt + byt * byt + bzt * bzt) + F(1.e-12)
    vcbx = (vvy * bzt - vvz * byt) / bmag
    vcby = -(vvx * bzt - vvz * bxt) / bmag
    vcbz = (vvx * byt - vvy * bxt) / bmag
    vcmag = np.sqrt(vcbx * vcbx + vcby * vcby + vcbz * vcbz) + F(1.e-12)
    ax, ay, az = bxt * x, byt * y, bzt * z
    p_para[b] = np.sqrt(ax * ax + ay * ay + az * az) / bmag
    cx, cy, cz = vcbx * x, vcby * y, vcbz * z
    p_cross[b] = np.sqrt(cx * cx + cy * cy + cz * cz) / vcmag
    pa, pc = p_para[b], p_cross[b]
    p_perp[b] = np.sqrt(np.abs(presmag * presmag - (pa * pa + pc * pc)))


# ---------------------------------------------------------------------------
# code/cft_findgrid.f90:4-60.  grid_minvals(3, n_grids) -> numpy
# (n_grids, 3); gridpt(4) -> int32 array (1-based indices and box).


def findgrid(xyz, n_grids, grid_minvals, grid_maxvals, xspac, gridpt, re_equiv, cname=None):
    abit = F(0.001)
    xa = (np.asarray(xyz, F) / F(re_equiv)).astype(F)
    gridpt[3] = 0
    for box in range(1, n_grids + 1):
        mn, mx = grid_minvals[box - 1], grid_maxvals[box - 1]
        if bool(np.all(xa > mn)) and bool(np.all(xa < mx)):
            gridpt[3] = box
            break
    if gridpt[3] == 0:
        gridpt[3] = n_grids
        xa = np.maximum(xa, grid_minvals[n_grids - 1] + abit)
        xa = np.minimum(xa, grid_maxvals[n_grids - 1] - abit)
    bb = int(gridpt[3]) - 1
    for axis in range(3):
        delta = F(xa[axis] - grid_minvals[bb, axis])
        gridpt[axis] = 1 + int(F(delta / xspac[bb]))


def aux_cube(gridpt, grid_minvals, xspac, re_equiv):
    """The cube_vertices(3,2) of code/cft_find_aux_cube.f90:62-65, as
    numpy (2, 3)."""
    bb = int(gridpt[3]) - 1
    cv = np.zeros((2, 3), F)
    for axis in range(3):
        cv[0, axis] = (grid_minvals[bb, axis] + (F(gridpt[axis]) - F(1.)) * xspac[bb]) \
            * F(re_equiv)
        cv[1, axis] = cv[0, axis] + xspac[bb] * F(re_equiv)
    return cv


# ---------------------------------------------------------------------------
# code/sub_flux_counter.f90:5-158


def flux_counter(qpx, qpy, qpz, hpx, hpy, hpz, opx, opy, opz, nx, ny, nz, n_grids, box, rmassq,
                 rmassh, rmasso, grid_spacing, qflux_in, qflux_out, hflux_in, hflux_out,
                 oflux_in, oflux_out):
    """Outputs are 1-element arrays.  Wall sums are sequential in the
    reference's loop order."""
    b = box - 1
    gs = F(grid_spacing)
    rm = (F(rmassq), F(rmassh), F(rmasso))
    inn = [F(0.)] * 3
    out = [F(0.)] * 3
    walls = (((qpx, hpx, opx), lambda a, i: a[b][:, :, i]),     # (k, j) rows: k outer, j inner
             ((qpy, hpy, opy), lambda a, j: a[b][:, j, :]),     # k outer, i inner
             ((qpz, hpz, opz), lambda a, k: a[b][k, :, :]))     # j outer, i inner
    ext = (nx, ny, nz)
    for axis, (arrs, cut) in enumerate(walls):
        f1 = [seq_sum(cut(a, 1).ravel()) for a in arrs]
        f2 = [seq_sum(cut(a, ext[axis] - 2).ravel()) for a in arrs]
        for s in range(3):
            if f1[1] > 0:
                inn[s] = inn[s] + f1[s] * gs / rm[s]
            else:
                out[s] = out[s] - f1[s] * gs / rm[s]
        for s in range(3):
            if f2[1] < 0:
                inn[s] = inn[s] - f2[s] * gs / rm[s]
            else:
                out[s] = out[s] + f2[s] * gs / rm[s]
    qflux_in[0], hflux_in[0], oflux_in[0] = inn
    qflux_out[0], hflux_out[0], oflux_out[0] = out


def linterp(xyzt1, xyzt2, time, xyz):
    """code/cft_linterp.f90:7-30: xyz (3,) at `time` on the straight line
    between the space-time points xyzt1 and xyzt2 (4,)."""
    a, b = np.asarray(xyzt1, F), np.asarray(xyzt2, F)
    if a[3] == b[3]:
        xyz[:] = a[:3]
    else:
        t = F(time)
        for ax in range(3):
            xyz[ax] = a[ax] + (b[ax] - a[ax]) / (b[3] - a[3]) * (t - a[3])


def craft_position(xcraftn, zcraftn, ut):
    """code/cft_find_aux_cube.f90:43-47: the craft's position at time ut."""
    s = np.zeros(3, F)
    if F(xcraftn[3]) == F(zcraftn[3]) or F(ut) >= F(zcraftn[3]):
        s[:] = np.asarray(xcraftn[:3], F)
    else:
        linterp(xcraftn, zcraftn, ut, s)
    return s


def find_aux_cube(cname, i_craft, scout, n_grids, grid_minvals, grid_maxvals, xspac, re_equiv,
                  xcraftn, zcraftn, ut, ntimesn, recordingn, gridpt, sxyz, cube_vertices):
    """code/cft_find_aux_cube.f90:4-67 (the unit close of a gridding problem
    is I/O and is left out); recordingn and cube_vertices are arrays (the
    latter (2, 3), i.e. the Fortran (3,2))."""
    sxyz[:] = craft_position(xcraftn, zcraftn, ut)
    findgrid(sxyz, n_grids, grid_minvals, grid_maxvals, xspac, gridpt, re_equiv)
    if gridpt[0] <= 0:
        recordingn[0] = 0
        return
    cube_vertices[...] = aux_cube(gridpt, grid_minvals, xspac, re_equiv)


def sample_craft(xcraft, zcraft, ut, n_grids, grid_minvals, grid_maxvals, xspac, re_equiv,
                 fields):
    """The aux-craft branch of craft_handling for many craft: positions at
    ut from (ncraft, 4) trajectory points, then sample_fields.  Returns
    (sxyz, scdata, gridpt)."""
    sxyz = np.stack([craft_position(x, z, ut) for x, z in zip(xcraft, zcraft)]).astype(F)
    sc, gp = sample_fields(sxyz, n_grids, grid_minvals, grid_maxvals, xspac, re_equiv, fields)
    return sxyz, sc, gp


def sample_fields(sxyz, n_grids, grid_minvals, grid_maxvals, xspac, re_equiv, fields):
    """Craft sampling (code/cft_craft_handling.f90:165-195 with
    code/cft_record_aux_data.f90:27-30): for each point, findgrid, the aux
    cube, the 2x2x2 cube of each field and trilin_interp.  ``sxyz`` is
    (npts, 3); fields are (n_grids, nz, ny, nx); returns (scdata (npts,
    nq), gridpt (npts, 4)).  A point whose cube would leave its box is
    marked gridpt[0] = 0 with zero samples (the reference's 'gridding
    problem' indicator)."""
    npts = sxyz.shape[0]
    nq = len(fields)
    out = np.zeros((npts, nq), F)
    gps = np.zeros((npts, 4), np.int32)
    for p in range(npts):
        gp = np.zeros(4, np.int32)
        findgrid(sxyz[p], n_grids, grid_minvals, grid_maxvals, xspac, gp, re_equiv)
        nz, ny, nx = fields[0].shape[1:]
        if not (1 <= gp[0] <= nx - 1 and 1 <= gp[1] <= ny - 1 and 1 <= gp[2] <= nz - 1):
            gp[0] = 0
            gps[p] = gp
            continue
        cv = aux_cube(gp, grid_minvals, xspac, re_equiv)
        i, j, k, bb = gp[0] - 1, gp[1] - 1, gp[2] - 1, gp[3] - 1
        for q in range(nq):
            cube = np.ascontiguousarray(fields[q][bb, k:k + 2, j:j + 2, i:i + 2])
            r = np.zeros(1, F)
            trilin_interp(np.asarray(sxyz[p], F), cv, cube, r)
            out[p, q] = r[0]
        gps[p] = gp
    return out, gps


def divB_correct_f64(bx, by, bz, rx, method="direct", rtol=1e-13):
    """All-FP64 divB_correct (code/sub_divB_correct.f90:130-172) of one box
    slice (nz, ny, nx): the central-difference right side, a solve of the
    same 7-band system in double -- exact (scipy's sparse direct solve,
    small grids) or by conjugate gradients to relative residual ``rtol``
    (method "cg"; the system is symmetric negative definite) -- then phi =
    rx^2 * x and B -= grad phi on the interior, all in double.  Returns the
    three corrected fields (float64)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    nz, ny, nx = bx.shape
    n = nx * ny * nz
    M = divb_matrix(nx, ny, nz).astype(np.float64)
    pos = [nx * ny, nx, 1, 0, -1, -nx, -nx * ny]
    rows, cols, vals = [], [], []
    for b in range(7):
        r = np.arange(n)
        c = r + pos[b]
        ok = (c >= 0) & (c < n)
        v = M[c[ok], b]
        keep = v != 0
        rows.append(r[ok][keep])
        cols.append(c[ok][keep])
        vals.append(v[keep])
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n, n))
    X, Y, Z = (a.astype(np.float64) for a in (bx, by, bz))
    rx = float(rx)
    rhs = np.zeros((nz, ny, nx))
    rhs[1:-1, 1:-1, 1:-1] = ((X[1:-1, 1:-1, 2:] - X[1:-1, 1:-1, :-2])
                             + (Y[1:-1, 2:, 1:-1] - Y[1:-1, :-2, 1:-1])
                             + (Z[2:, 1:-1, 1:-1] - Z[:-2, 1:-1, 1:-1])) / (2.0 * rx)
    if method == "direct":
        x = spl.spsolve(A.tocsc(), rhs.ravel())
    else:
        x, info = spl.cg(-A, -rhs.ravel(), rtol=rtol, atol=0.0, maxiter=20 * (nx + ny + nz))
        if info != 0:
            raise RuntimeError(f"FP64 CG did not converge ({info})")
    phi = (x * rx * rx).reshape(nz, ny, nx)
    out = [X.copy(), Y.copy(), Z.copy()]
    out[0][1:-1, 1:-1, 1:-1] -= (phi[1:-1, 1:-1, 2:] - phi[1:-1, 1:-1, :-2]) / (2.0 * rx)
    out[1][1:-1, 1:-1, 1:-1] -= (phi[1:-1, 2:, 1:-1] - phi[1:-1, :-2, 1:-1]) / (2.0 * rx)
    out[2][1:-1, 1:-1, 1:-1] -= (phi[2:, 1:-1, 1:-1] - phi[:-2, 1:-1, 1:-1]) / (2.0 * rx)
    return out


# ---------------------------------------------------------------------------
# Initialisation: the dipole field (code/sub_dipole.f90), mak_dip_grd
# (code/sub_mak_dip_grd.f90), set_resist (code/sub_set_resist.f90) and the
# main program's initialisation block (code/3d_main.f90:1238-1633).
# sin / cos / pow of single-precision values are the C library's sinf,
# cosf and powf, which a compiled reference calls; integer powers x**n are
# expanded left to right, ((x*x)*x)*..., as flang expands them.

import ctypes as _ct
import ctypes.util as _ctu

_LIBM = _ct.CDLL(_ctu.find_library("m"))
for _fn in ("sinf", "cosf"):
    getattr(_LIBM, _fn).restype = _ct.c_float
    getattr(_LIBM, _fn).argtypes = [_ct.c_float]
_LIBM.powf.restype = _ct.c_float
_LIBM.powf.argtypes = [_ct.c_float, _ct.c_float]


def _sinf(x):
    return F(_LIBM.sinf(float(F(x))))


def _cosf(x):
    return F(_LIBM.cosf(float(F(x))))


# How the initialisation's real power and the torus's exponentials are
# evaluated: "libm" calls the C library's powf / expf, as the compiled
# reference does; "double" evaluates them in double precision and rounds
# once to single, as the device kernels do.  The two differ in the last bit
# for a small fraction of arguments; everything else is unaffected.
TRANSCENDENTALS = "libm"


def set_transcendentals(kind: str) -> None:
    global TRANSCENDENTALS
    if kind not in ("libm", "double"):
        raise ValueError(kind)
    TRANSCENDENTALS = kind


def _powf(a, e):
    """Elementwise powf of a float32 array (small grids only)."""
    a = np.asarray(a, F)
    if TRANSCENDENTALS == "double":
        return np.power(a.astype(np.float64), float(F(e))).astype(F)
    out = np.fromiter((_LIBM.powf(float(v), float(F(e))) for v in a.ravel()), F, a.size)
    return out.reshape(a.shape)


def _ipow(x, n):
    r = x
    for _ in range(n - 1):
        r = r * x
    return r


def _dipole_core(x1, y1, z1, sin_rot, cos_rot, sin_tilt, cos_tilt, b0):
    """code/sub_dipole.f90:17-49 (and the same lines of mak_dip_grd); returns
    (abx, aby, abz, ar)."""
    xr = x1 * cos_rot + y1 * sin_rot
    yr = -(x1 * sin_rot) + y1 * cos_rot
    xp = xr * cos_tilt - z1 * sin_tilt
    yp = yr
    zp = xr * sin_tilt + z1 * cos_tilt
    x2, y2, z2 = xp * xp, yp * yp, zp * zp
    ar = np.sqrt(x2 + y2 + z2)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        bmag = -(b0 / _ipow(ar, 5))
        dbx = -(F(3.) * bmag * xp * zp)
        dby = -(F(3.) * bmag * yp * zp)
        dbz = bmag * (x2 + y2 - F(2.) * z2)
        rbx = dbx * cos_tilt + dbz * sin_tilt
        rby = dby
        rbz = -(dbx * sin_tilt) + dbz * cos_tilt
        abx = rbx * cos_rot - rby * sin_rot
        aby = rbx * sin_rot + rby * cos_rot
    return abx, aby, rbz, ar


def dipole(abx, aby, abz, ax, ay, az, rot_angle=0.0, xdip=0.0, ydip=0.0, zdip=0.0,
           sin_tilt=0.0, cos_tilt=1.0, b0=1.0):
    """code/sub_dipole.f90:4-52 at one point; abx/aby/abz are 1-element
    arrays; the keyword arguments are common /rotation/."""
    x1 = F(ax) - F(xdip)
    y1 = F(ay) - F(ydip)
    z1 = F(az) - F(zdip)
    bx, by, bz, _ = _dipole_core(x1, y1, z1, _sinf(rot_angle), _cosf(rot_angle), F(sin_tilt),
                                 F(cos_tilt), F(b0))
    abx[0], aby[0], abz[0] = bx, by, bz


def _box_axes(nx, ny, nz, x0, x1, y0, y1, z0, z1):
    dx = (F(x1) - F(x0)) / (F(nx) - F(1.))
    dy = (F(y1) - F(y0)) / (F(ny) - F(1.))
    dz = (F(z1) - F(z0)) / (F(nz) - F(1.))
    ax = (F(x0) + dx * np.arange(nx, dtype=F))[None, None, :]
    ay = (F(y0) + dy * np.arange(ny, dtype=F))[None, :, None]
    az = (F(z0) + dz * np.arange(nz, dtype=F))[:, None, None]
    return ax, ay, az


def tilt_rotation(old_tilt, utold, t_old_box, dt_off, dtilt, t_equiv, planet_per):
    """The dipole's tilt and the planet's rotation angle at t_old + dt_off:
    the half step before step 1 (code/3d_main.f90:2173-2180) or the full step
    before step 2 (:2465-2472), in the reference's mixed precision (t_old,
    utold double; everything else single; pi = 4*atan(1.0) in single).
    Returns (sin_tilt, cos_tilt, rot_angle) as float32."""
    pi = F(np.pi)   # 4*atan(1.0) in single precision, 3.14159274
    tb = float(t_old_box) + float(F(dt_off))
    atilt = F(float(F(old_tilt)) + tb * float(F(dtilt)))
    arg = F(F(atilt * pi) / F(180.0))
    sin_tilt, cos_tilt = _sinf(arg), _cosf(arg)
    ut2 = F(float(utold) + tb * float(F(t_equiv)) / 3600.0)
    if F(planet_per) == 0:   # no rotation (the reference divides by 0 here)
        return sin_tilt, cos_tilt, F(0.0)
    nrot2 = int(F(ut2 / F(planet_per)))
    rot_hr2 = F(ut2 - F(F(nrot2) * F(planet_per)))
    rot_angle = F(F(F(F(2.0) * pi) * rot_hr2) / F(planet_per))
    return sin_tilt, cos_tilt, rot_angle


def main_clock(t, ut, tilt, delt, t_equiv, dtilt):
    """code/3d_main.f90:1966-1977: the full step's clock advance; returns
    (t, ut, tilt, utold, old_tilt) -- t and ut double, tilt single."""
    utold, old_tilt = ut, F(tilt)
    t = t + float(F(delt))
    ut = t * float(F(t_equiv)) / 3600.0
    tilt = F(F(tilt) + F(F(dtilt) * F(delt)))
    return t, ut, tilt, utold, old_tilt


def mak_dip_grd(bx0, by0, bz0, nx, ny, nz, n_grids, mbndry, box, ijzero, numzero, mzero,
                r_inner, grd_xmin, grd_xmax, grd_ymin, grd_ymax, grd_zmin, grd_zmax,
                rot_angle=0.0, xdip=0.0, ydip=0.0, zdip=0.0, sin_tilt=0.0, cos_tilt=1.0,
                b0=1.0):
    """code/sub_mak_dip_grd.f90:4-112; the keyword arguments are common
    /rotation/."""
    b = box - 1
    ax, ay, az = _box_axes(nx, ny, nz, grd_xmin[b], grd_xmax[b], grd_ymin[b], grd_ymax[b],
                           grd_zmin[b], grd_zmax[b])
    x1, y1, z1 = ax - F(xdip), ay - F(ydip), az - F(zdip)
    bx, by, bz, ar = _dipole_core(x1, y1, z1, _sinf(rot_angle), _cosf(rot_angle),
                                  F(sin_tilt), F(cos_tilt), F(b0))
    keep = ar > F(r_inner) - F(1.5)
    bx0[b] = np.where(keep, bx, F(0.))
    by0[b] = np.where(keep, by, F(0.))
    bz0[b] = np.where(keep, bz, F(0.))
    if box <= mbndry:
        for n in range(int(numzero[b])):
            i, j, k = (int(ijzero[n, c, b]) - 1 for c in range(3))
            bx0[b, k, j, i] = by0[b, k, j, i] = bz0[b, k, j, i] = F(0.)


def set_resist(rst, nx, ny, nz, mbndry, resist, ijzero, numzero, ijmid, nummid, ijsrf, numsrf,
               msrf, mmid, mzero, b0):
    """code/sub_set_resist.f90:10-63 (the counts of box mbndry for every
    box, as the reference takes them)."""
    rst[...] = F(0.)
    mb = mbndry
    for lst, num, val in ((ijzero, numzero, F(b0) / F(resist)),
                          (ijmid, nummid, F(0.5) * F(b0) / F(resist)),
                          (ijsrf, numsrf, F(0.125) * F(b0) / F(resist))):
        for n in range(int(num[mb - 1])):
            for b in range(mb):
                i, j, k = (int(lst[n, c, b]) - 1 for c in range(3))
                rst[b, k, j, i] = val


def init_plasma(f, bx0, by0, bz0, resistive, nx, ny, nz, n_grids, mbndry, msrf, mmid, mzero,
                grid_minvals, grid_maxvals, parm_srf, parm_mid, parm_zero, ijsrf, ijmid, ijzero,
                numsrf, nummid, numzero, s):
    """code/3d_main.f90:1238-1633, the main program's initialisation block.
    ``f`` maps the 34 state-field names to (n_grids, nz, ny, nx) arrays;
    ``s`` holds the scalars by their reference names (rmassq ... b0, with
    rot_angle and the common /rotation/ values).  Lists and counts are
    filled in place.  Returns the list lengths found (before any cut)."""
    g = {k: F(v) for k, v in s.items()}
    sin_rot, cos_rot = _sinf(s["rot_angle"]), _cosf(s["rot_angle"])
    st, ct = g["sin_tilt"], g["cos_tilt"]
    numsrf[...] = 0
    nummid[...] = 0
    numzero[...] = 0
    found = np.zeros((mbndry, 3), np.int64)
    ri = g["r_inner"]
    for b in range(n_grids):
        ax, ay, az = _box_axes(nx, ny, nz, grid_minvals[b, 0], grid_maxvals[b, 0],
                               grid_minvals[b, 1], grid_maxvals[b, 1], grid_minvals[b, 2],
                               grid_maxvals[b, 2])
        az, ay, ax = az - g["zdip"], ay - g["ydip"], ax - g["xdip"]
        shp = (nz, ny, nx)
        ax, ay, az = (np.broadcast_to(a, shp).astype(F) for a in (ax, ay, az))
        xr = ax * cos_rot + ay * sin_rot
        yr = -(ax * sin_rot) + ay * cos_rot
        xp = xr * ct - az * st
        yp = yr
        zp = xr * st + az * ct
        ar = np.sqrt(xp * xp + yp * yp + zp * zp)
        bx, by, bz, _ = _dipole_core(ax - g["xdip"], ay - g["ydip"], az - g["zdip"], sin_rot,
                                     cos_rot, st, ct, g["b0"])
        inner = ar < g["zero_bndry"]
        bx0[b] = np.where(inner, F(0.), bx)
        by0[b] = np.where(inner, F(0.), by)
        bz0[b] = np.where(inner, F(0.), bz)
        rx = ax * g["re_equiv"]
        ry = ay * g["re_equiv"]
        rd = np.sqrt(rx * rx + ry * ry)
        rr = g["r_rot"]
        t = (F(2.) * (rr * rr)) / (rr * rr + rd * rd)
        vfrac = np.where(rd < rr, F(1.), t * t)
        rvy = rx * g["v_rot"] * vfrac
        rvx = -(ry * g["v_rot"] * vfrac)
        ar_iono = np.sqrt(xp * xp + ay * ay + F(1.25) * (zp * zp))
        ar_iono = np.maximum(F(0.0001), ar_iono)
        ra = _powf((ar_iono + F(0.5) * ri) / (F(1.5) * ri), -g["alpha_e"])
        h15, h30 = F(1.5) * ri, F(3.0) * ri
        zheight = np.maximum(F(1.), (zp * zp + h15 * h15) / (h30 * h30))
        ra = ra / (zheight * zheight)
        rho_iono = np.maximum(g["erho"] * ra, g["d_min"])
        arho = np.maximum(rho_iono, g["d_min"])
        vt = np.maximum(np.sqrt(rvy * rvy + rvx * rvx), g["cs_inner"])
        qrho = arho * g["rmassq"] / zheight
        qpres = arho * vt
        hrho = F(0.05) * qrho * g["rmassh"] / g["rmassq"]
        hpres = F(0.05) * qpres
        orho = qrho * g["rmasso"] / g["rmassq"]
        for sp, rho, pr in (("q", qrho, qpres), ("h", hrho, hpres), ("o", orho, qpres)):
            f[sp + "rho"][b] = rho
            for c in "xyz":
                f[sp + "pres" + c][b] = pr
            for c in ("xy", "xz", "yz"):
                f[sp + "pres" + c][b] = F(0.)
            f[sp + "px"][b] = rho * rvx
            f[sp + "py"][b] = rho * rvy
            f[sp + "pz"][b] = F(0.)
        f["epres"][b] = (qpres + hpres + qpres) / g["ti_te"]
        for c in "xyz":
            f["b" + c][b] = F(0.)
        if b < mbndry:
            # code/3d_main.f90:1389-1452: loop order k, j, i = memory order
            cls = np.zeros(shp, np.int8)
            near = ar <= g["srf_bndry"]
            cls[near] = 1
            cls[near & (ar < g["mid_bndry"])] = 2
            cls[near & (ar < g["zero_bndry"])] = 3
            flat = cls.ravel()
            vals = [f[k][b].ravel() for k in ("qrho", "hrho", "orho", "qpresx", "hpresx",
                                              "opresx", "epres")]
            for q, (ij, parm, cap) in enumerate(((ijsrf, parm_srf, msrf), (ijmid, parm_mid, mmid),
                                                 (ijzero, parm_zero, mzero))):
                idx = np.nonzero(flat == q + 1)[0]
                found[b, q] = len(idx)
                idx = idx[:cap]
                kk, jj, ii = np.unravel_index(idx, shp)
                ij[:len(idx), 0, b] = ii + 1
                ij[:len(idx), 1, b] = jj + 1
                ij[:len(idx), 2, b] = kk + 1
                for a in range(7):
                    parm[:len(idx), a, b] = vals[a][idx]
                (numsrf, nummid, numzero)[q][b] = len(idx)
    # solar wind (:1464-1587) and the speed check (:1589-1626)
    wb = g["r_rot"] / g["re_equiv"]
    ofrac = g["rho_frac"]
    for b in range(n_grids):
        ax, ay, az = _box_axes(nx, ny, nz, grid_minvals[b, 0], grid_maxvals[b, 0],
                               grid_minvals[b, 1], grid_maxvals[b, 1], grid_minvals[b, 2],
                               grid_maxvals[b, 2])
        az, ay, ax = az - g["zdip"], ay - g["ydip"], ax - g["xdip"]
        shp = (nz, ny, nx)
        ax, ay, az = (np.broadcast_to(a, shp).astype(F) for a in (ax, ay, az))
        ar = np.sqrt(ax * ax + ay * ay + az * az)
        full = (ar >= F(1.5) * wb) & (ax < F(0.))
        ramp = (ar > wb) & (ar < F(1.5) * wb) & (ax < F(0.))
        d = ((ar - wb) / (F(0.5) * wb)).astype(F)

        def sc(v, m):   # an added term, times dfrac on the ramp
            return np.where(m, v * d, v) if np.ndim(v) else np.where(m, F(v) * d, F(v))

        for m, on in ((full, full), (ramp, ramp)):
            if not on.any():
                continue
            rm = m & ramp

            def add(key, term):
                f[key][b] = np.where(m, sc(term, rm) + f[key][b], f[key][b])

            add("qrho", g["srho"])
            add("qpx", g["spx"])
            add("qpy", g["spy"])
            add("qpz", g["spz"])
            qr = f["qrho"][b]
            with np.errstate(divide="ignore", invalid="ignore"):
                avx, avy, avz = f["qpx"][b] / qr, f["qpy"][b] / qr, f["qpz"][b] / qr
            apres = g["cs_wind"] * g["cs_wind"] * qr / g["gamma"]
            add("qpresx", F(0.5) * apres)
            add("epres", F(0.5) * apres / g["ti_te"])
            add("hrho", g["srho"] * g["rho_frac"] * g["rmassh"] / g["rmassq"])
            add("hpresx", F(0.5) * g["spress"] * g["rho_frac"])
            add("orho", g["srho"] * ofrac * g["rmasso"] / g["rmassq"])
            add("opresx", F(0.5) * g["spress"] * ofrac)
            for sp in "qho":
                for c in "yz":
                    f[sp + "pres" + c][b] = np.where(m, f[sp + "presx"][b],
                                                     f[sp + "pres" + c][b])
                for c in ("xy", "xz", "yz"):
                    f[sp + "pres" + c][b] = np.where(m, F(0.), f[sp + "pres" + c][b])
            for sp in "ho":
                for c, av in zip("xyz", (avx, avy, avz)):
                    f[sp + "p" + c][b] = np.where(m, av * f[sp + "rho"][b], f[sp + "p" + c][b])
        for sp in "qho":
            r = f[sp + "rho"][b]
            with np.errstate(divide="ignore", invalid="ignore"):
                vx, vy, vz = f[sp + "px"][b] / r, f[sp + "py"][b] / r, f[sp + "pz"][b] / r
                spd = np.sqrt(vx * vx + vy * vy + vz * vz)
            fast = spd > F(1.)
            for c in "xyz":
                f[sp + "p" + c][b] = np.where(fast, F(0.), f[sp + "p" + c][b])
    set_resist(resistive, nx, ny, nz, mbndry, s["resist"], ijzero, numzero, ijmid, nummid,
               ijsrf, numsrf, msrf, mmid, mzero, F(1.))
    return found


# ---------------------------------------------------------------------------
# code/3d_main.f90:3149-3253, the torus source (inline in the main program)

_LIBM.expf.restype = _ct.c_float
_LIBM.expf.argtypes = [_ct.c_float]


def _expf(a):
    a = np.asarray(a, F)
    if TRANSCENDENTALS == "double":
        return np.exp(a.astype(np.float64)).astype(F)
    out = np.fromiter((_LIBM.expf(float(v)) for v in a.ravel()), F, a.size)
    return out.reshape(a.shape)


def torus_inject(f, bx0, by0, bz0, nx, ny, nz, grid_minvals, grid_maxvals, s):
    """Adds the torus plasma to box 1 of the fields ``f`` (names -> (n_grids,
    nz, ny, nx) arrays, updated in place); ``s`` holds the driver's scalars
    by name.  Returns (tot_q, tot_h, tot_o), sum of den*dx*dy*dz in the
    reference's loop order."""
    g = {k: F(v) for k, v in s.items()}
    x0, x1 = grid_minvals[0], grid_maxvals[0]
    ax, ay, az = _box_axes(nx, ny, nz, x0[0], x1[0], x0[1], x1[1], x0[2], x1[2])
    dx = (F(x1[0]) - F(x0[0])) / (F(nx) - F(1.))
    dy = (F(x1[1]) - F(x0[1])) / (F(ny) - F(1.))
    dz = (F(x1[2]) - F(x0[2])) / (F(nz) - F(1.))
    shp = (nz, ny, nx)
    az, ay, ax = az - g["zdip"], ay - g["ydip"], ax - g["xdip"]
    ax, ay, az = (np.broadcast_to(a, shp).astype(F) for a in (ax, ay, az))
    rx, ry = ax * g["re_equiv"], ay * g["re_equiv"]
    ar = np.sqrt((ax + g["xdip"]) * (ax + g["xdip"]) + (ay + g["ydip"]) * (ay + g["ydip"])
                 + (az + g["zdip"]) * (az + g["zdip"]))
    rvy = rx * g["v_rot"]
    rvx = -(ry * g["v_rot"])
    corotate = np.sqrt(rvx * rvx + rvy * rvy)
    ar_tmid = np.sqrt(ax * ax + ay * ay) * g["re_equiv"]
    dr_tmid = np.abs(ar_tmid - g["torus_inj_dist"])
    m = (np.abs(dr_tmid) < F(2.) * g["torus_rad"]) & (ar > g["srf_bndry"])
    a1 = dr_tmid / (F(0.5) * g["torus_rad"])
    a2 = (az * g["re_equiv"]) / (F(0.5) * g["torus_rad"])
    rscale = np.where(m, _expf(-(a1 * a1)), F(0.))
    zscale = np.where(m, _expf(-(a2 * a2)), F(0.))
    bsx, bsy, bsz = bx0[0] + f["bx"][0], by0[0] + f["by"][0], bz0[0] + f["bz"][0]
    with np.errstate(divide="ignore", invalid="ignore"):
        abtot = (bsx * bsx + bsy * bsy) / (bsz * bsz)
    dscale = np.maximum(F(0.), F(1.) - F(10.) * abtot)
    c2 = corotate * corotate
    dels, dens = {}, {}
    for sp, dn, rm in (("h", "denh_torus", "rmassh"), ("q", "denq_torus", "rmassq"),
                       ("o", "deno_torus", "rmasso")):
        den = g[dn] * g[rm] * rscale * zscale * dscale
        dp = den * c2 * g["t_torus"] if sp == "h" else (den / g[rm]) * c2 * g["t_torus"]
        for key, add in (("rho", den), ("presx", dp), ("presy", dp),
                         ("presz", dp * g["aniso_factor"]), ("presxy", dp),
                         ("px", g["reduct"] * den * rvx), ("py", g["reduct"] * den * rvy)):
            f[sp + key][0] = np.where(m, f[sp + key][0] + add, f[sp + key][0])
        dels[sp], dens[sp] = dp, den
    f["epres"][0] = np.where(m, f["epres"][0] + (dels["o"] + dels["h"] + dels["q"]),
                             f["epres"][0])
    tots = []
    for sp in "qho":
        contrib = np.where(m, dens[sp] * dx * dy * dz, F(0.))
        tots.append(seq_sum(contrib.ravel()))
    return tuple(tots)
