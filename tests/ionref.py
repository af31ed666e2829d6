"""Test infrastructure: torch restatements of push_ion (one species) and
lap_plasma, written for torch.autograd, dtype-generic.

Float32 runs reproduce the oracle's restatements (oracle/restate.py
push_ion, lap_plasma) bit for bit, which pins them; float64 runs give the
vector-Jacobian products the device adjoints (g3m3_push_ion_adjoint,
g3m3_lap_plasma_adjoint, g3m3_species_velocity_adjoint) are held to.

Fields are box-local (nz, ny, nx) tensors.  Species order of the ten fields:
rho, presx, presy, presz, px, py, pz, presxy, presxz, presyz (the library's
IonPtrs order).  Reference: code/sub_push_ion.f90:4-362,
code/sub_lap_plasma.f90:4-147, code/3d_main.f90:2698-2703 (velocities).
"""
from __future__ import annotations

import numpy as np
import torch

FIELDS10 = ("rho", "presx", "presy", "presz", "px", "py", "pz", "presxy", "presxz", "presyz")
RHO, PX, PY, PZ = 0, 4, 5, 6
PRX, PRY, PRZ, PXY, PXZ, PYZ = 1, 2, 3, 7, 8, 9


def _i(a, dz=0, dy=0, dx=0):
    nz, ny, nx = a.shape
    return a[1 + dz:nz - 1 + dz, 1 + dy:ny - 1 + dy, 1 + dx:nx - 1 + dx]


def _sqrt(x):
    """Correctly rounded square root (torch's float32 CPU sqrt is not, in a
    few cells in a thousand): float32 through float64, which rounds once."""
    return torch.sqrt(x.double()).to(x.dtype) if x.dtype == torch.float32 else torch.sqrt(x)


def _set(a, v):
    a = a.clone()
    a[1:-1, 1:-1, 1:-1] = v
    return a


def push_ion(out_init, old, src, bs, btot, efld, tv, sc, dtype):
    """One species: returns the ten output fields (interior replaced,
    boundary = out_init's).  sc: rmass, delt, grav, re_equiv, reynolds,
    grd (xmin, xmax, ymin, ymax, zmin, zmax) -- float32 values."""
    S = lambda x: torch.tensor(float(np.float32(x)), dtype=dtype)
    nz, ny, nx = src[0].shape
    g = sc["grd"]
    ddx = (S(g[1]) - S(g[0])) / S(nx - 1)
    ddy = (S(g[3]) - S(g[2])) / S(ny - 1)
    ddz = (S(g[5]) - S(g[4])) / S(nz - 1)
    two = S(2)
    dxt, dyt, dzt = two * ddx, two * ddy, two * ddz
    rmass, delt, grav = S(sc["rmass"]), S(sc["delt"]), S(sc["grav"])
    re_equiv, reynolds = S(sc["re_equiv"]), S(sc["reynolds"])
    bmax = S(0.20) * rmass / delt / reynolds
    wr, wx, wy, wz, wpx, wpy, wpz, wxy, wxz, wyz = src
    aq = torch.clamp(wr, min=float(np.float32(1e-6)))
    qvx, qvy, qvz = wpx / aq, wpy / aq, wpz / aq

    def d(a, ax):
        if ax == 0:
            return _i(a, 0, 0, 1) - _i(a, 0, 0, -1)
        if ax == 1:
            return _i(a, 0, 1) - _i(a, 0, -1)
        return _i(a, 1) - _i(a, -1)

    def flux(w):
        return ((d(w * qvx, 0) / dxt) + (d(w * qvy, 1) / dyt)) + (d(w * qvz, 2) / dzt)

    C = _i
    rho = C(old[RHO]) - delt * (((d(wpx, 0) / dxt) + (d(wpy, 1) / dyt)) + (d(wpz, 2) / dzt))
    aqpx = C(old[PX]) - delt * flux(wpx)
    aqpy = C(old[PY]) - delt * flux(wpy)
    aqpz = C(old[PZ]) - delt * flux(wpz)
    abx, aby, abz = C(bs[0]), C(bs[1]), C(bs[2])
    bmag = C(btot)
    weight = torch.clamp(bmag / bmax, min=1.0)
    skin = reynolds / weight
    tx, ty, tz = C(tv[0]), C(tv[1]), C(tv[2])
    vx, vy, vz = C(qvx), C(qvy), C(qvz)
    tcx = ty * abz - tz * aby
    tcy = -(tx * abz - tz * abx)
    tcz = tx * aby - ty * abx
    qcx = vy * abz - vz * aby
    qcy = -(vx * abz - vz * abx)
    qcz = vx * aby - vy * abx
    gpx = d(wx, 0) / dxt + d(wxy, 1) / dxt + d(wxz, 2) / dxt
    gpy = d(wxy, 0) / dxt + d(wy, 1) / dyt + d(wyz, 2) / dxt
    gpz = d(wxz, 0) / dxt + d(wyz, 1) / dyt + d(wz, 2) / dzt
    dex = (C(efld[0]) + tcx) * reynolds
    dey = (C(efld[1]) + tcy) * reynolds
    dez = (C(efld[2]) + tcz) * reynolds
    dvx = (qcx - tcx) * skin
    dvy = (qcy - tcy) * skin
    dvz = (qcz - tcz) * skin
    qden = C(wr) / rmass
    px_ = aqpx + delt * (qden * (dex + dvx) - gpx)
    py_ = aqpy + delt * (qden * (dey + dvy) - gpy)
    pz_ = aqpz + delt * (qden * (dez + dvz) - gpz)
    ax = (S(g[0]) + ddx * torch.arange(1, nx - 1, dtype=dtype))[None, None, :]
    ay = (S(g[2]) + ddy * torch.arange(1, ny - 1, dtype=dtype))[None, :, None]
    az = (S(g[4]) + ddz * torch.arange(1, nz - 1, dtype=dtype))[:, None, None]
    if dtype == torch.float32:     # the reference's float32 abscissae
        ax = (S(g[0]) + ddx * torch.arange(1, nx - 1).to(dtype))[None, None, :]
    radius = _sqrt(ax * ax + ay * ay + az * az) + S(0.0000001)
    t = re_equiv * radius
    gg = delt * C(wr) * grav / (t * t)
    npx = px_ - gg * ax / radius
    npy = py_ - gg * ay / radius
    npz = pz_ - gg * az / radius
    px0 = C(old[PRX]) - delt * flux(wx) - two * delt * C(wx) * d(qvx, 0) / dxt
    py0 = C(old[PRY]) - delt * flux(wy) - two * delt * C(wy) * d(qvy, 1) / dyt
    pz0 = C(old[PRZ]) - delt * flux(wz) - two * delt * C(wz) * d(qvz, 2) / dzt

    def sm(b):
        return (S(6) * C(b) + _i(b, 0, 0, 1) + _i(b, 0, 0, -1) + _i(b, 0, 1) + _i(b, 0, -1)
                + _i(b, 1) + _i(b, -1)) / S(12)
    sbx, sby, sbz = sm(bs[0]), sm(bs[1]), sm(bs[2])
    bm2 = _sqrt(sbx * sbx + sby * sby + sbz * sbz)
    sf = torch.where(bm2 > bmax, reynolds * (bmax / torch.where(bm2 > bmax, bm2, bmax)),
                     reynolds * torch.ones_like(bm2))
    cxy, cxz, cyz = C(wxy), C(wxz), C(wyz)
    cx, cy, cz = C(wx), C(wy), C(wz)
    npresx = (px0 - two * delt * cxy * d(qvx, 1) / dyt - two * delt * cxz * d(qvx, 2) / dzt
              + two * delt * sf / rmass * (cxy * sbz - cxz * sby))
    npresy = (py0 - two * delt * cxy * d(qvy, 0) / dxt - two * delt * cyz * d(qvy, 2) / dzt
              + two * delt * sf / rmass * (cyz * sbx - cxy * sbz))
    npresz = (pz0 - two * delt * cxz * d(qvz, 0) / dxt - two * delt * cyz * d(qvz, 1) / dyt
              + two * delt * sf / rmass * (cxz * sby - cyz * sbx))
    pxy = C(old[PXY]) - delt * flux(wxy)
    npxy = (pxy - delt * cxy * (d(qvx, 0) / dxt + d(qvy, 1) / dyt)
            - delt * cy * d(qvx, 1) / dyt - delt * cx * d(qvy, 0) / dxt
            - delt * cyz * d(qvx, 2) / dzt - delt * cxz * d(qvy, 2) / dzt
            + delt * sf / rmass * (sbz * (cy - cx) - sby * cyz + sbx * cxz))
    pxz = C(old[PXZ]) - delt * flux(wxz)
    npxz = (pxz - delt * cxz * (d(qvx, 0) / dxt + d(qvz, 2) / dzt)
            - delt * cz * d(qvx, 2) / dzt - delt * cx * d(qvz, 0) / dxt
            - delt * cyz * d(qvx, 1) / dyt - delt * cxy * d(qvz, 1) / dyt
            + delt * sf / rmass * (sby * (cx - cz) - sbx * cxy + sbz * cyz))
    pyz = C(old[PYZ]) - delt * flux(wyz)
    npyz = (pyz - delt * cyz * (d(qvy, 1) / dyt + d(qvz, 2) / dzt)
            - delt * cy * d(qvz, 1) / dyt - delt * cz * d(qvy, 2) / dzt
            - delt * cxy * d(qvz, 0) / dxt - delt * cxz * d(qvy, 0) / dxt
            + delt * sf / rmass * (sbx * (cz - cy) - sbz * cxz + sby * cxy))
    vals = [rho, npresx, npresy, npresz, npx, npy, npz, npxy, npxz, npyz]
    return [_set(o, v) for o, v in zip(out_init, vals)]


def lap_plasma(q, w_init, v, chirho, chipxyz, delt, dtype):
    """The ten fields of one species smoothed with velocity v (chirho for
    rho, chipxyz for the rest); boundary = w_init's."""
    S = lambda x: torch.tensor(float(np.float32(x)), dtype=dtype)
    dl = S(delt)
    vx, vy, vz = v
    uxp = torch.abs(_i(vx, 0, 0, 1) - _i(vx))
    uxm = torch.abs(_i(vx) - _i(vx, 0, 0, -1))
    uyp = torch.abs(_i(vy, 0, 1) - _i(vy))
    uym = torch.abs(_i(vy) - _i(vy, 0, -1))
    uzp = torch.abs(_i(vz, 1) - _i(vz))
    uzm = torch.abs(_i(vz) - _i(vz, -1))
    out = []
    for f, (qq, w) in enumerate(zip(q, w_init)):
        chi = S(chirho if f == RHO else chipxyz)
        c = _i(qq)
        val = c + chi * (dl * (uxp * (_i(qq, 0, 0, 1) - c) - uxm * (c - _i(qq, 0, 0, -1)))
                         + dl * (uyp * (_i(qq, 0, 1) - c) - uym * (c - _i(qq, 0, -1)))
                         + dl * (uzp * (_i(qq, 1) - c) - uzm * (c - _i(qq, -1))))
        out.append(_set(w, val))
    return out


def species_velocity(rho, p, smallbit, dtype):
    """v = p / max(rho, smallbit) (code/3d_main.f90:2698-2703)."""
    den = torch.clamp(rho, min=float(np.float32(smallbit)))
    return [p[0] / den, p[1] / den, p[2] / den]
