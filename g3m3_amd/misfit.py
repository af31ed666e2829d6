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
two forward runs per parameter, stepped together on several streams (a
3 x 81x81x41 run alone leaves most of the GPU idle).  That is the check an
adjoint would be held to, and its cost is what an adjoint (at most 4 forward
runs) would replace.
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


def craft_positions(cfg: MisfitConfig, wl) -> np.ndarray:
    """(ncraft, 3) positions in planet radii, the unit findgrid takes
    (code/cft_findgrid.f90 divides by re_equiv to reach grid units): r ~
    U[r_min, r_max], direction uniform on the sphere.  A draw that does not
    lie inside the outer box at least two of its cells from every face is
    drawn again, so every craft samples the field inside a box."""
    g = wl.grid
    b = cfg.ngrd - 1
    re = float(wl.scal["re_equiv"])
    margin = 2.0 * float(g["xspac"][b])
    lo = np.array([g["grd_xmin"][b], g["grd_ymin"][b], g["grd_zmin"][b]], np.float64) + margin
    hi = np.array([g["grd_xmax"][b], g["grd_ymax"][b], g["grd_zmax"][b]], np.float64) - margin
    rng = np.random.default_rng(cfg.seed)
    out = []
    for _ in range(10000 * cfg.ncraft):
        if len(out) == cfg.ncraft:
            break
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        p = d * rng.uniform(cfg.r_min, cfg.r_max)
        u = p / re
        if np.all(u >= lo) and np.all(u <= hi):
            out.append(p)
    if len(out) < cfg.ncraft:
        raise ValueError("no craft position in [r_min, r_max] fits inside the outer box")
    return np.ascontiguousarray(np.array(out).astype(F))


def workload_for(theta, cfg: MisfitConfig):
    t = [float(x) for x in theta]
    return make_workload("nested121", L=cfg.L, ngrd=cfg.ngrd, imf=(t[0], t[1], t[2]),
                         n_sw=t[3], v_sw=t[4], eta=t[5])


class _Sampler:
    """The craft samples of one run: b x, y, z and the ion density."""

    NAMES = ("bx", "by", "bz", "qrho", "hrho", "orho")

    def __init__(self, run, wl, cfg: MisfitConfig):
        import torch
        self.run, self.wl, self.cfg = run, wl, cfg
        self.pos = torch.from_numpy(craft_positions(cfg, wl)).cuda()
        nq = len(self.NAMES)
        self.out = torch.zeros((cfg.ncraft, nq), dtype=torch.float32, device="cuda")
        self.gp = torch.zeros((cfg.ncraft, 4), dtype=torch.int32, device="cuda")
        self.ptrs = (_ct.c_void_p * nq)(*[run.a[k].data_ptr() for k in self.NAMES])
        g = wl.grid
        self.mn = np.ascontiguousarray(np.stack([g["grd_xmin"], g["grd_ymin"], g["grd_zmin"]],
                                                1).astype(F))
        self.mx = np.ascontiguousarray(np.stack([g["grd_xmax"], g["grd_ymax"], g["grd_zmax"]],
                                                1).astype(F))
        self.xs = np.ascontiguousarray(g["xspac"], dtype=F)
        self.samples = []

    def sample(self):
        wl, cfg = self.wl, self.cfg
        self.run.sync_arrays()     # the state back in the reference's arrays
        _lib.call("sample_fields", cfg.ncraft, self.pos, cfg.ngrd, self.mn, self.mx, self.xs,
                  F(wl.scal["re_equiv"]), wl.nx, wl.ny, wl.nz, len(self.NAMES), self.ptrs,
                  self.out, self.gp)
        h = self.out.cpu().numpy()
        self.samples.append(np.stack([h[:, 0], h[:, 1], h[:, 2],
                                      (h[:, 3] + h[:, 4]) + h[:, 5]], 1).astype(F))


def forward(theta, cfg: MisfitConfig = MisfitConfig()):
    """Samples (steps // every, ncraft, 4): b x, y, z and the ion density at
    each craft after every ``every`` full steps; and the device time of the
    steps and samples (seconds, host set-up excluded)."""
    return forward_many([theta], cfg, streams=1)


def forward_many(thetas, cfg: MisfitConfig = MisfitConfig(), streams: int = 8):
    """forward() of several parameter vectors at once: the runs step
    together on ``streams`` streams (each run's results equal the run alone,
    bit for bit).  Returns the list of samples (a single array for one run)
    and the device time of all of them."""
    import torch
    from .driver import Batch
    wls = [workload_for(t, cfg) for t in thetas]
    batch = Batch(wls, streams=streams)
    batch.init_dt()
    samplers = [_Sampler(r, wl, cfg) for r, wl in zip(batch.runs, wls)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for st in range(1, cfg.steps + 1):
        batch.full_step()
        if st % cfg.every == 0:
            for sm in samplers:
                sm.sample()
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    _lib.check("misfit forward")
    out = [np.stack(sm.samples) for sm in samplers]
    return (out[0] if len(out) == 1 else out), secs


def loss(samples, target) -> float:
    """sum((b - b*)^2 + (rho - rho*)^2) over craft and sampling times, in
    double precision."""
    d = samples.astype(np.float64) - target.astype(np.float64)
    return float(np.sum(d * d))


def gradient_fd(theta, target, cfg: MisfitConfig = MisfitConfig(), rel: float = 1e-3,
                together: bool = True, samples_out: list | None = None):
    """Central-difference gradient of the loss with respect to theta, step
    rel * |theta_i| (rel where theta_i = 0), and the device seconds its
    forward runs took: all 2 len(theta) runs stepped together, or (together
    False) one after another.  Both give the same gradient, bit for bit.
    samples_out: a list that receives the runs' samples (theta_0 + h, theta_0
    - h, theta_1 + h, ...)."""
    theta = np.asarray(theta, dtype=np.float64)
    hs, runs = [], []
    for i in range(len(theta)):
        h = rel * abs(theta[i]) if theta[i] != 0.0 else rel
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        hs.append(h)
        runs += [tp, tm]
    if together:
        samples, secs = forward_many(runs, cfg)
    else:
        samples, secs = [], 0.0
        for t in runs:
            sm, a = forward(t, cfg)
            samples.append(sm)
            secs += a
    g = np.array([(loss(samples[2 * i], target) - loss(samples[2 * i + 1], target)) /
                  (2.0 * hs[i]) for i in range(len(theta))])
    if samples_out is not None:
        samples_out.extend(samples)
    return g, secs
