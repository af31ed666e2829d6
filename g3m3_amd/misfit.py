"""The differentiation configuration's forward model, loss and gradient.

Three nested 81x81x41 boxes advance 50 full steps.  Every 10 steps, eight
spacecraft sample the perturbation field b and the ion density (the three
species' rho summed) at their positions, r uniform in [8, 20] planet radii
and the direction uniform on the sphere.  Sampling is the reference's
findgrid + aux cube + trilin_interp (code/cft_craft_handling.f90:165-195), run
on the device by ``g3m3_sample_fields``.  The loss is the squared misfit
sum((b - b*)^2 + (rho - rho*)^2) against a target run.

The gradient with respect to theta = (IMF x, y, z, upstream n, upstream v,
uniform resistivity) is formed by central differences at h = 1e-3 |theta_i|:
two forward runs per parameter.  That is the check an adjoint would be held
to, and its cost is what an adjoint (at most 4 forward runs) would replace.
"""
from __future__ import annotations

import ctypes as _ct
import time
from dataclasses import dataclass

import numpy as np

from . import _lib
from .workload import make_workload

F = np.float32

# theta: IMF (nT) x, y, z; upstream density (cm^-3); upstream speed (km/s);
# uniform resistivity (normalised)
THETA_NAMES = ("imf_x", "imf_y", "imf_z", "n_sw", "v_sw", "eta")
THETA0 = np.array([0.0, 0.0, -5.0, 5.0, 400.0, 1e-3])


@dataclass
class MisfitConfig:
    L: int = 40               # 81 x 81 x 41 boxes
    ngrd: int = 3
    steps: int = 50
    every: int = 10           # sampling interval (full steps)
    ncraft: int = 8
    seed: int = 20261206      # craft positions
    r_min: float = 8.0        # planet radii
    r_max: float = 20.0


def craft_positions(cfg: MisfitConfig, re_equiv: float) -> np.ndarray:
    """(ncraft, 3) positions in grid units: r ~ U[r_min, r_max] planet radii,
    direction uniform on the sphere (a planet radius is 1/re_equiv units)."""
    rng = np.random.default_rng(cfg.seed)
    d = rng.normal(size=(cfg.ncraft, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = rng.uniform(cfg.r_min, cfg.r_max, cfg.ncraft) / re_equiv
    return np.ascontiguousarray((d * r[:, None]).astype(F))


def workload_for(theta, cfg: MisfitConfig):
    t = [float(x) for x in theta]
    return make_workload("nested121", L=cfg.L, ngrd=cfg.ngrd, imf=(t[0], t[1], t[2]),
                         n_sw=t[3], v_sw=t[4], eta=t[5])


def forward(theta, cfg: MisfitConfig = MisfitConfig()):
    """Samples (steps // every, ncraft, 4): b x, y, z and the ion density at
    each craft after every ``every`` full steps; and the device time of the
    steps and samples (seconds, host set-up excluded)."""
    import torch
    from .driver import FastRun
    wl = workload_for(theta, cfg)
    run = FastRun(wl)
    run.init_dt()
    pos = craft_positions(cfg, float(wl.scal["re_equiv"]))
    dpos = torch.from_numpy(pos).cuda()
    nq = 6
    out = torch.zeros((cfg.ncraft, nq), dtype=torch.float32, device="cuda")
    gp = torch.zeros((cfg.ncraft, 4), dtype=torch.int32, device="cuda")
    names = ("bx", "by", "bz", "qrho", "hrho", "orho")
    ptrs = (_ct.c_void_p * nq)(*[run.a[k].data_ptr() for k in names])
    mn = np.ascontiguousarray(np.stack([wl.grid["grd_xmin"], wl.grid["grd_ymin"],
                                        wl.grid["grd_zmin"]], 1).astype(F))
    mx = np.ascontiguousarray(np.stack([wl.grid["grd_xmax"], wl.grid["grd_ymax"],
                                        wl.grid["grd_zmax"]], 1).astype(F))
    xs = np.ascontiguousarray(wl.grid["xspac"], dtype=F)
    samples = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for st in range(1, cfg.steps + 1):
        run.full_step()
        if st % cfg.every == 0:
            run.sync_arrays()     # the state back in the reference's arrays
            _lib.call("sample_fields", cfg.ncraft, dpos, cfg.ngrd, mn, mx, xs,
                      F(wl.scal["re_equiv"]), wl.nx, wl.ny, wl.nz, nq, ptrs, out, gp)
            h = out.cpu().numpy()
            samples.append(np.stack([h[:, 0], h[:, 1], h[:, 2],
                                     (h[:, 3] + h[:, 4]) + h[:, 5]], 1).astype(F))
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    _lib.check("misfit forward")
    return np.stack(samples), secs


def loss(samples, target) -> float:
    """sum((b - b*)^2 + (rho - rho*)^2) over craft and sampling times, in
    double precision."""
    d = samples.astype(np.float64) - target.astype(np.float64)
    return float(np.sum(d * d))


def gradient_fd(theta, target, cfg: MisfitConfig = MisfitConfig(), rel: float = 1e-3):
    """Central-difference gradient of the loss with respect to theta, step
    rel * |theta_i| (rel where theta_i = 0); the device seconds of all the
    forward runs it took."""
    theta = np.asarray(theta, dtype=np.float64)
    g = np.zeros_like(theta)
    secs = 0.0
    for i in range(len(theta)):
        h = rel * abs(theta[i]) if theta[i] != 0.0 else rel
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        sp, a = forward(tp, cfg)
        sm, b = forward(tm, cfg)
        secs += a + b
        g[i] = (loss(sp, target) - loss(sm, target)) / (2.0 * h)
    return g, secs
