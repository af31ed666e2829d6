"""Gradients of a spacecraft misfit through the magnetic-field update, by the
discrete adjoint.

The model is the field half of the reference's box-step on one outermost box
with the plasma held fixed (``g3m3_field_model_*`` in include/g3m3.h): two
Lax-Wendroff stages of calcur / total B / fnd_evel / BandE (+ a uniform
resistivity eta) / push_bfld / bndry_outer, then lap_bfld, fcsmooth and
bndry_outer again (code/3d_main.f90:2243-2945 for box = n_grids = 1).  The
controls are ``ctl = (eta, sbx_wind, sby_wind, sbz_wind)``: the resistivity
and the upstream field that bndry_outer imposes on the inflow face.  The
solar-wind density and speed enter the reference's step through the plasma,
which this model holds fixed, so they are not controls here.

Virtual spacecraft sample b (the perturbation field, as the reference's
craft do, code/cft_craft_handling.f90:178-189) every ``every`` steps through
``g3m3_sample_fields`` (findgrid + trilin_interp).  ``FieldRollout`` is a
``torch.autograd.Function`` from (ctl, b at step 0) to the samples: the
forward pass keeps b every ``ckpt`` steps, and when a gradient is wanted
the tapes (what the adjoint of a step needs) of the last segment; the
backward pass recomputes every earlier segment from its checkpoint with
tapes and runs the hand-written adjoint kernels back through each
(``g3m3_field_model_adjoint``, ``g3m3_sample_fields_adjoint``).
"""
from __future__ import annotations

import ctypes as _ct
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .workload import B_EQUIV_NT, EV, P_EQUIV_PA, RHO_EQUIV_CC, V_EQUIV_KMS

F = np.float32
PLASMA = ("qrho", "hrho", "orho", "qpx", "qpy", "qpz", "hpx", "hpy", "hpz",
          "opx", "opy", "opz", "epres")
CTL_NAMES = ("eta", "sbx_wind", "sby_wind", "sbz_wind")


class FieldModelDesc(_ct.Structure):
    """Mirror of g3m3_field_model_desc (include/g3m3.h)."""
    _fields_ = [("nx", _ct.c_int), ("ny", _ct.c_int), ("nz", _ct.c_int),
                ("rho", _ct.c_void_p * 3), ("mom", _ct.c_void_p * 9), ("epres", _ct.c_void_p),
                ("b0", _ct.c_void_p * 3)] + [
        (k, _ct.c_float) for k in ("rx", "delt", "rmassq", "rmassh", "rmasso", "reynolds",
                                   "smallbit", "chierg", "diferg")]


@dataclass
class FieldSetup:
    """One box of the model, host arrays shaped (nz, ny, nx) float32."""
    nx: int
    ny: int
    nz: int
    plasma: dict                 # PLASMA -> array
    b0: np.ndarray               # (3, nz, ny, nx)
    b_init: np.ndarray           # (3, nz, ny, nx)
    ctl: np.ndarray              # (4,) eta, sbx, sby, sbz
    scal: dict                   # rx, delt, rmass*, reynolds, smallbit, chierg, diferg
    grid: dict = field(default_factory=dict)   # grid_minvals/maxvals (1,3), xspac (1,)


def _modes(rng, shape, xs, n=3):
    """A smooth seeded field in [-1, 1]: the mean of n random plane waves."""
    nz, ny, nx = shape
    x = (np.arange(nx) * xs)[None, None, :]
    y = (np.arange(ny) * xs)[None, :, None]
    z = (np.arange(nz) * xs)[:, None, None]
    out = np.zeros(shape)
    ext = max(nx, ny, nz) * xs
    for _ in range(n):
        k = rng.normal(size=3) * (2.0 * math.pi / ext) * rng.uniform(1.0, 3.0)
        out += np.cos(k[0] * x + k[1] * y + k[2] * z + rng.uniform(0, 2 * math.pi))
    return out / n


def dipole_b0(ax, ay, az) -> np.ndarray:
    """(3, nz, ny, nx) float32: the dipole of make_workload's planet
    (|B| = 10 b_equiv on the equator at r = 3) at the float32 node
    coordinates ax (1, 1, nx), ay (1, ny, 1), az (nz, 1, 1)."""
    ar = np.sqrt(ax * ax + ay * ay + az * az).astype(F)
    b0m = F(10.0 * 27.0)
    bmag = -b0m / ar ** 5
    return np.stack([-3.0 * bmag * ax * az, -3.0 * bmag * ay * az,
                     bmag * (ax * ax + ay * ay - 2.0 * az * az)]).astype(F)


def field_setup(nx: int = 41, ny: int = 41, nz: int = 21, spacing: float = 0.5,
                seed: int = 20261207, n_sw: float = 5.0, v_sw: float = 400.0,
                t_sw: float = 10.0, imf=(0.0, 0.0, -5.0), eta: float = 1e-3,
                standoff: float = 6.0) -> FieldSetup:
    """A box upstream of a dipole planet in the reference's normalised units
    (the conventions of g3m3_amd.workload): solar wind along +x with smooth
    seeded +-10% variations of density and flow, the dipole of
    make_workload's planet (|B| = 10 b_equiv on the equator at r = 3), the
    box's downstream face ``standoff`` grid units from the planet.  b starts
    at the upstream field minus the dipole plus +-5% smooth variations.
    Lapidus and flux-correction coefficients are make_input's defaults
    (chierg = 1, diferg = 0.0025; scripts/make_input.f90:196-199)."""
    rng = np.random.default_rng(seed)
    h = F(spacing)
    shape = (nz, ny, nx)
    xmin = F(-standoff - (nx - 1) * float(h))
    ymin = F(-(ny - 1) * float(h) / 2.0)
    zmin = F(-(nz - 1) * float(h) / 2.0)
    ax = (xmin + h * np.arange(nx, dtype=F))[None, None, :]
    ay = (ymin + h * np.arange(ny, dtype=F))[None, :, None]
    az = (zmin + h * np.arange(nz, dtype=F))[:, None, None]
    b0 = dipole_b0(ax, ay, az)
    srho = n_sw / RHO_EQUIV_CC
    svel = v_sw / V_EQUIV_KMS
    spress = n_sw * 1e6 * t_sw * EV / P_EQUIV_PA
    dens = srho * (1.0 + 0.1 * _modes(rng, shape, float(h)))
    vx = svel * (1.0 + 0.1 * _modes(rng, shape, float(h)))
    vy = svel * 0.1 * _modes(rng, shape, float(h))
    vz = svel * 0.1 * _modes(rng, shape, float(h))
    pres = 0.5 * spress * (1.0 + 0.1 * _modes(rng, shape, float(h)))
    rho_frac, o_conc, rmassq, rmassh, rmasso = 0.001, 0.125, 1.0, 1.0, 16.0
    pl = {}
    for s, frac in (("q", 1.0), ("h", rho_frac * rmassh / rmassq),
                    ("o", rho_frac * o_conc * rmasso / rmassq)):
        r = (dens * frac).astype(F)
        pl[s + "rho"] = r
        pl[s + "px"] = (r * vx).astype(F)
        pl[s + "py"] = (r * vy).astype(F)
        pl[s + "pz"] = (r * vz).astype(F)
    pl["epres"] = (pres * (1.0 + rho_frac + rho_frac * o_conc)).astype(F)
    sb = np.array(imf, np.float64) / B_EQUIV_NT
    b_init = np.stack([(sb[c] - b0[c]) * (1.0 + 0.05 * _modes(rng, shape, float(h)))
                       for c in range(3)]).astype(F)
    # time step: a quarter cell per step at the fastest of flow + Alfven speed
    btot = np.sqrt(((b_init + b0) ** 2).sum(0))
    vmax = float(np.max(np.sqrt(vx ** 2 + vy ** 2 + vz ** 2) + btot / np.sqrt(dens)))
    scal = dict(rx=h, delt=F(0.25 * float(h) / vmax), rmassq=F(rmassq), rmassh=F(rmassh),
                rmasso=F(rmasso), reynolds=F(256.0), smallbit=F(1e-5), chierg=F(1.0),
                diferg=F(0.0025))
    grid = dict(grid_minvals=np.array([[xmin, ymin, zmin]], F),
                grid_maxvals=np.array([[xmin + h * (nx - 1), ymin + h * (ny - 1),
                                        zmin + h * (nz - 1)]], F),
                xspac=np.array([h], F), re_equiv=F(1.0))
    ctl = np.array([eta, sb[0], sb[1], sb[2]], F)
    return FieldSetup(nx=nx, ny=ny, nz=nz, plasma=pl, b0=b0, b_init=b_init, ctl=ctl,
                      scal=scal, grid=grid)


def craft_positions(setup: FieldSetup, n: int, seed: int = 20261208, margin: float = 2.0):
    """(n, 3) float32 positions (grid units = planet radii here, re_equiv
    = 1) uniform in the box at least ``margin`` cells from every face."""
    g = setup.grid
    h = float(g["xspac"][0])
    lo = g["grid_minvals"][0].astype(np.float64) + margin * h
    hi = g["grid_maxvals"][0].astype(np.float64) - margin * h
    rng = np.random.default_rng(seed)
    return np.ascontiguousarray(rng.uniform(lo, hi, size=(n, 3)).astype(F))


class FieldModel:
    """The device model of one FieldSetup (g3m3_field_model_create)."""

    def __init__(self, setup: FieldSetup):
        import torch
        self.setup = setup
        self.nx, self.ny, self.nz = setup.nx, setup.ny, setup.nz
        self.V = setup.nx * setup.ny * setup.nz
        self.t = {k: torch.from_numpy(np.ascontiguousarray(setup.plasma[k])).cuda()
                  for k in PLASMA}
        self.b0 = torch.from_numpy(np.ascontiguousarray(setup.b0)).cuda()
        d = FieldModelDesc()
        d.nx, d.ny, d.nz = setup.nx, setup.ny, setup.nz
        for i, k in enumerate(("qrho", "hrho", "orho")):
            d.rho[i] = self.t[k].data_ptr()
        for i, k in enumerate(PLASMA[3:12]):
            d.mom[i] = self.t[k].data_ptr()
        d.epres = self.t["epres"].data_ptr()
        for c in range(3):
            d.b0[c] = self.b0[c].data_ptr()
        for k, v in setup.scal.items():
            setattr(d, k, float(v))
        lib = _lib.lib()
        lib.g3m3_field_model_create.restype = _ct.c_void_p
        lib.g3m3_field_model_tape_floats.restype = _ct.c_longlong
        lib.g3m3_field_model_tape_floats.argtypes = [_ct.c_void_p]
        for fn in ("step", "adjoint"):
            getattr(lib, "g3m3_field_model_" + fn).restype = _ct.c_int
        self.h = lib.g3m3_field_model_create(_ct.byref(d))
        if not self.h:
            raise _lib.G3M3Error("g3m3_field_model_create failed")
        self.tape_floats = int(lib.g3m3_field_model_tape_floats(_ct.c_void_p(self.h)))
        self._lib = lib
        g = setup.grid
        self._grid = (np.ascontiguousarray(g["grid_minvals"]), np.ascontiguousarray(g["grid_maxvals"]),
                      np.ascontiguousarray(g["xspac"]), F(g["re_equiv"]))

    def close(self):
        if self.h:
            self._lib.g3m3_field_model_destroy(_ct.c_void_p(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _ctl(ctl) -> np.ndarray:
        return np.ascontiguousarray(np.asarray(ctl, dtype=F).reshape(4))

    def new_state(self, b=None):
        import torch
        src = self.setup.b_init if b is None else b
        return torch.from_numpy(np.ascontiguousarray(src, dtype=F)).cuda().reshape(3, self.V)

    def new_tape(self):
        import torch
        return torch.empty(self.tape_floats, dtype=torch.float32, device="cuda")

    def step(self, b_in, b_out, ctl, tape=None):
        c = self._ctl(ctl)
        rc = self._lib.g3m3_field_model_step(
            _ct.c_void_p(self.h), _ct.c_void_p(b_in.data_ptr()), _ct.c_void_p(b_out.data_ptr()),
            _ct.c_void_p(tape.data_ptr() if tape is not None else 0),
            c.ctypes.data_as(_ct.c_void_p))
        if rc != 0:
            raise _lib.G3M3Error(f"g3m3_field_model_step returned {rc}")

    def adjoint(self, tape, bout_bar, bin_bar, ctl_bar, ctl):
        c = self._ctl(ctl)
        rc = self._lib.g3m3_field_model_adjoint(
            _ct.c_void_p(self.h), _ct.c_void_p(tape.data_ptr()),
            _ct.c_void_p(bout_bar.data_ptr()), _ct.c_void_p(bin_bar.data_ptr()),
            _ct.c_void_p(ctl_bar.data_ptr()), c.ctypes.data_as(_ct.c_void_p))
        if rc != 0:
            raise _lib.G3M3Error(f"g3m3_field_model_adjoint returned {rc}")

    def _ptrs(self, b):
        return (_ct.c_void_p * 3)(*[b[c].data_ptr() for c in range(3)])

    def sample(self, b, pos, out, gp):
        mn, mx, xs, re = self._grid
        _lib.call("sample_fields", pos.shape[0], pos, 1, mn, mx, xs, re, self.nx, self.ny,
                  self.nz, 3, self._ptrs(b), out, gp)

    def sample_adjoint(self, pos, s_bar, b_bar):
        mn, mx, xs, re = self._grid
        _lib.call("sample_fields_adjoint", pos.shape[0], pos, 1, mn, mx, xs, re, self.nx,
                  self.ny, self.nz, 3, self._ptrs(b_bar), s_bar)


def _function():
    import torch

    class FieldRollout(torch.autograd.Function):
        """(ctl, b at step 0) -> samples of b after every ``every``-th of
        ``steps`` steps, shape (steps // every, npts, 3).  ctl is a float64
        CPU tensor of 4 (rounded to float32 for the device); b a (3, V)
        float32 device tensor.  The backward pass returns the gradient in
        ctl (float64, accumulated in double on the device) and in b."""

        @staticmethod
        def forward(ctx, ctl, b_start, model, steps, every, pos, ckpt, keep):
            c = ctl.detach().cpu().numpy()
            ns = steps // every
            out = torch.empty((ns, pos.shape[0], 3), dtype=torch.float32, device="cuda")
            gp = torch.empty((pos.shape[0], 4), dtype=torch.int32, device="cuda")
            b = b_start.detach().clone()
            nb = torch.empty_like(b)
            ckpts = []
            last = (steps - 1) // ckpt * ckpt
            tapes = [model.new_tape() for _ in range(last, steps)] if keep else None
            for t in range(steps):
                if t % ckpt == 0:
                    ckpts.append(b.clone())
                model.step(b, nb, c, tapes[t - last] if keep and t >= last else None)
                b, nb = nb, b
                if (t + 1) % every == 0:
                    model.sample(b, pos, out[(t + 1) // every - 1], gp)
            ctx.model, ctx.steps, ctx.every, ctx.ckpt = model, steps, every, ckpt
            ctx.c, ctx.pos, ctx.ckpts, ctx.last_tapes = c, pos, ckpts, tapes
            ctx.ctl_meta = (ctl.dtype, ctl.device, ctl.shape)
            return out

        @staticmethod
        def backward(ctx, gs):
            model, steps, every, ckpt = ctx.model, ctx.steps, ctx.every, ctx.ckpt
            gs = gs.contiguous()
            b_bar = torch.zeros((3, model.V), dtype=torch.float32, device="cuda")
            tmp = torch.empty_like(b_bar)
            ctl_bar = torch.zeros(4, dtype=torch.float64, device="cuda")
            for seg in reversed(range(0, steps, ckpt)):
                end = min(seg + ckpt, steps)
                tapes = ctx.last_tapes
                ctx.last_tapes = None
                if tapes is None:
                    # recompute the segment from its checkpoint (the last
                    # segment's tapes come from the forward pass)
                    tapes = [model.new_tape() for _ in range(seg, end)]
                    b = ctx.ckpts[seg // ckpt].clone()
                    nb = torch.empty_like(b)
                    for t in range(seg, end):
                        model.step(b, nb, ctx.c, tapes[t - seg])
                        b, nb = nb, b
                for t in reversed(range(seg, end)):
                    if (t + 1) % every == 0:
                        model.sample_adjoint(ctx.pos, gs[(t + 1) // every - 1], b_bar)
                    model.adjoint(tapes[t - seg], b_bar, tmp, ctl_bar, ctx.c)
                    b_bar, tmp = tmp, b_bar
                del tapes
            dtype, device, shape = ctx.ctl_meta
            g_ctl = ctl_bar.cpu().to(dtype).reshape(shape).to(device)
            return g_ctl, b_bar, None, None, None, None, None, None

    return FieldRollout


_FR = None


def rollout(ctl, b_start, model: FieldModel, steps: int, every: int, pos, ckpt: int = 10):
    """Samples of the model's rollout, differentiable in ctl and b_start."""
    import torch
    global _FR
    if _FR is None:
        _FR = _function()
    keep = torch.is_grad_enabled() and (ctl.requires_grad or b_start.requires_grad)
    return _FR.apply(ctl, b_start, model, steps, every, pos, ckpt, keep)


# --- the whole box-step -------------------------------------------------------
#
# ``g3m3_full_model_*`` (include/g3m3.h): one box-step of an outermost box
# with every field evolving -- the ion pushes, push_elec, push_bfld,
# Lapidus and flux correction of code/3d_main.f90:2163-3030 -- and its
# adjoint.  The state is the 34 fields in the canonical order of
# g3m3_amd.workload.FIELDS plus the electron pressure of the last Lapidus
# pass (the reference's wrkepres, which the next step's push_elec and
# stage-2 BandE read).  The controls are the resistivity and the nine
# upstream values bndry_outer imposes on the inflow face.

FULL_CTL_NAMES = ("eta", "sbx_wind", "sby_wind", "sbz_wind", "srho", "spx", "spy", "spz",
                  "spress")
NSTATE = 35
#: state indices of the quantities the full model's craft sample
FULL_SAMPLED = (0, 4, 5, 6, 30, 31, 32, 33)     # qrho, qpx, qpy, qpz, epres, bx, by, bz


class FullModelDesc(_ct.Structure):
    """Mirror of g3m3_full_model_desc (include/g3m3.h)."""
    _fields_ = [("nx", _ct.c_int), ("ny", _ct.c_int), ("nz", _ct.c_int),
                ("b0", _ct.c_void_p * 3)] + [
        (k, _ct.c_float) for k in ("rx", "xmin", "xmax", "ymin", "ymax", "zmin", "zmax",
                                   "delt", "rmassq", "rmassh", "rmasso", "reynolds",
                                   "smallbit", "gamma", "gamma1", "grav", "re_equiv",
                                   "chirho", "chipxyz", "chierg", "difrho", "diferg",
                                   "rho_frac", "o_conc", "ti_te")]


@dataclass
class FullSetup:
    """One outermost box with all its fields, host arrays float32."""
    nx: int
    ny: int
    nz: int
    state: np.ndarray            # (35, nz, ny, nx)
    b0: np.ndarray               # (3, nz, ny, nx)
    ctl: np.ndarray              # (9,) FULL_CTL_NAMES
    scal: dict                   # the float fields of FullModelDesc
    grid: dict = field(default_factory=dict)   # grid_minvals/maxvals (1,3), xspac (1,)


def full_setup(nx: int = 41, ny: int = 41, nz: int = 21, spacing: float = 0.5,
               seed: int = 20261209, n_sw: float = 5.0, v_sw: float = 400.0,
               t_sw: float = 10.0, imf=(0.0, 0.0, -5.0), eta: float = 1e-3,
               standoff: float = 6.0, aniso: float = 0.05) -> FullSetup:
    """The box of field_setup with every state field: each species' density
    and flow as there (the h and o fractions of the reference's solar wind,
    code/sub_bndry_outer.f90:50-103), diagonal pressures with +-10% smooth
    variations, off-diagonal pressures of +-``aniso`` of the diagonal, the
    electron pressure the ions' sum over ti_te, and the Lapidus electron
    pressure of a previous step within +-2% of it.  The time step is a
    quarter cell at the fastest of flow + Alfven + sound speed."""
    fs = field_setup(nx, ny, nz, spacing, seed, n_sw, v_sw, t_sw, imf, eta, standoff)
    rng = np.random.default_rng(seed + 1)
    shape = (nz, ny, nx)
    h = float(fs.scal["rx"])
    rho_frac, o_conc, ti_te = 0.001, 0.125, 1.0
    spress = n_sw * 1e6 * t_sw * EV / P_EQUIV_PA
    pl = fs.plasma
    st = np.zeros((NSTATE,) + shape, F)
    pres_frac = {"q": 1.0, "h": rho_frac, "o": rho_frac * o_conc}
    ptot = np.zeros(shape)
    for si, s in enumerate("qho"):
        base = 10 * si
        st[base + 0] = pl[s + "rho"]
        p = 0.5 * spress * pres_frac[s] * (1.0 + 0.1 * _modes(rng, shape, h))
        for m in (1, 2, 3):
            st[base + m] = p * (1.0 + 0.02 * _modes(rng, shape, h))
        st[base + 4], st[base + 5], st[base + 6] = pl[s + "px"], pl[s + "py"], pl[s + "pz"]
        for m in (7, 8, 9):
            st[base + m] = aniso * p * _modes(rng, shape, h)
        ptot += st[base + 1]
    st[30] = (ptot / ti_te).astype(F)
    st[31:34] = fs.b_init
    st[34] = (st[30] * (1.0 + 0.02 * _modes(rng, shape, h))).astype(F)
    dens = st[0].astype(np.float64)
    v = np.sqrt(st[4].astype(np.float64) ** 2 + st[5] ** 2 + st[6] ** 2) / dens
    btot = np.sqrt(((fs.b_init + fs.b0) ** 2).sum(0))
    cs = np.sqrt(1.6666 * 2.0 * ptot / dens)
    vmax = float(np.max(v + np.sqrt(btot ** 2 / dens + cs ** 2)))
    g = fs.grid
    mn, mx = g["grid_minvals"][0], g["grid_maxvals"][0]
    sc = fs.scal
    scal = dict(rx=sc["rx"], xmin=mn[0], xmax=mx[0], ymin=mn[1], ymax=mx[1], zmin=mn[2],
                zmax=mx[2], delt=F(0.25 * h / vmax), rmassq=sc["rmassq"], rmassh=sc["rmassh"],
                rmasso=sc["rmasso"], reynolds=sc["reynolds"], smallbit=sc["smallbit"],
                gamma=F(1.6666), gamma1=F(0.6666), grav=F(0.0), re_equiv=F(1.0),
                chirho=F(2.0), chipxyz=F(2.0), chierg=F(1.0), difrho=F(0.01),
                diferg=F(0.0025), rho_frac=F(rho_frac), o_conc=F(o_conc), ti_te=F(ti_te))
    srho = n_sw / RHO_EQUIV_CC
    svel = v_sw / V_EQUIV_KMS
    ctl = np.array([eta, fs.ctl[1], fs.ctl[2], fs.ctl[3], srho, srho * svel, 0.0, 0.0,
                    spress], F)
    return FullSetup(nx=nx, ny=ny, nz=nz, state=st, b0=fs.b0, ctl=ctl, scal=scal, grid=g)


class FullModel:
    """The device model of one FullSetup (g3m3_full_model_create)."""

    def __init__(self, setup: FullSetup):
        import torch
        self.setup = setup
        self.nx, self.ny, self.nz = setup.nx, setup.ny, setup.nz
        self.V = setup.nx * setup.ny * setup.nz
        self.b0 = torch.from_numpy(np.ascontiguousarray(setup.b0)).cuda()
        d = FullModelDesc()
        d.nx, d.ny, d.nz = setup.nx, setup.ny, setup.nz
        for c in range(3):
            d.b0[c] = self.b0[c].data_ptr()
        for k, v in setup.scal.items():
            setattr(d, k, float(v))
        lib = _lib.lib()
        lib.g3m3_full_model_create.restype = _ct.c_void_p
        lib.g3m3_full_model_tape_floats.restype = _ct.c_longlong
        lib.g3m3_full_model_tape_floats.argtypes = [_ct.c_void_p]
        for fn in ("step", "adjoint"):
            getattr(lib, "g3m3_full_model_" + fn).restype = _ct.c_int
        self.h = lib.g3m3_full_model_create(_ct.byref(d))
        if not self.h:
            raise _lib.G3M3Error("g3m3_full_model_create failed")
        self.tape_floats = int(lib.g3m3_full_model_tape_floats(_ct.c_void_p(self.h)))
        self._lib = lib
        g = setup.grid
        self._grid = (np.ascontiguousarray(g["grid_minvals"]), np.ascontiguousarray(g["grid_maxvals"]),
                      np.ascontiguousarray(g["xspac"]), F(g["re_equiv"]))

    def close(self):
        if self.h:
            self._lib.g3m3_full_model_destroy(_ct.c_void_p(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _ctl(ctl) -> np.ndarray:
        return np.ascontiguousarray(np.asarray(ctl, dtype=F).reshape(len(FULL_CTL_NAMES)))

    def new_state(self, st=None):
        import torch
        src = self.setup.state if st is None else st
        return torch.from_numpy(np.ascontiguousarray(src, dtype=F)).cuda().reshape(NSTATE, self.V)

    def new_tape(self):
        import torch
        return torch.empty(self.tape_floats, dtype=torch.float32, device="cuda")

    def step(self, s_in, s_out, ctl, tape=None):
        c = self._ctl(ctl)
        rc = self._lib.g3m3_full_model_step(
            _ct.c_void_p(self.h), _ct.c_void_p(s_in.data_ptr()), _ct.c_void_p(s_out.data_ptr()),
            _ct.c_void_p(tape.data_ptr() if tape is not None else 0),
            c.ctypes.data_as(_ct.c_void_p))
        if rc != 0:
            raise _lib.G3M3Error(f"g3m3_full_model_step returned {rc}")

    def adjoint(self, tape, sout_bar, sin_bar, ctl_bar, ctl):
        c = self._ctl(ctl)
        rc = self._lib.g3m3_full_model_adjoint(
            _ct.c_void_p(self.h), _ct.c_void_p(tape.data_ptr()),
            _ct.c_void_p(sout_bar.data_ptr()), _ct.c_void_p(sin_bar.data_ptr()),
            _ct.c_void_p(ctl_bar.data_ptr()), c.ctypes.data_as(_ct.c_void_p))
        if rc != 0:
            raise _lib.G3M3Error(f"g3m3_full_model_adjoint returned {rc}")

    # --- as the inner box of a nested pair (g3m3_full_model_set_parent) ---
    def set_parent(self, parent: FullSetup):
        """Bound this box by flanks from ``parent`` (same extents in cells,
        a whole multiple of this box's spacing) and restrict it into the
        parent's core."""
        geom = np.array([self.setup.scal[k] for k in ("xmin", "xmax", "ymin", "ymax", "zmin",
                                                     "zmax")]
                        + [parent.scal[k] for k in ("xmin", "xmax", "ymin", "ymax", "zmin",
                                                    "zmax")], F)
        rc = self._lib.g3m3_full_model_set_parent(_ct.c_void_p(self.h),
                                                  geom.ctypes.data_as(_ct.c_void_p))
        if rc != 0:
            raise _lib.G3M3Error(f"g3m3_full_model_set_parent returned {rc}")

    def step_inner(self, s_in, s_out, ctl, p_old, p_cur, t12, tape=None):
        c = self._ctl(ctl)
        t = np.ascontiguousarray(np.asarray(t12, F).reshape(4))
        rc = self._lib.g3m3_full_model_step_inner(
            _ct.c_void_p(self.h), _ct.c_void_p(s_in.data_ptr()), _ct.c_void_p(s_out.data_ptr()),
            _ct.c_void_p(tape.data_ptr() if tape is not None else 0),
            c.ctypes.data_as(_ct.c_void_p), _ct.c_void_p(p_old.data_ptr()),
            _ct.c_void_p(p_cur.data_ptr()), t.ctypes.data_as(_ct.c_void_p))
        if rc != 0:
            raise _lib.G3M3Error(f"g3m3_full_model_step_inner returned {rc}")

    def adjoint_inner(self, tape, sout_bar, sin_bar, ctl_bar, ctl, t12, p_old_bar, p_cur_bar):
        c = self._ctl(ctl)
        t = np.ascontiguousarray(np.asarray(t12, F).reshape(4))
        rc = self._lib.g3m3_full_model_adjoint_inner(
            _ct.c_void_p(self.h), _ct.c_void_p(tape.data_ptr()),
            _ct.c_void_p(sout_bar.data_ptr()), _ct.c_void_p(sin_bar.data_ptr()),
            _ct.c_void_p(ctl_bar.data_ptr()), c.ctypes.data_as(_ct.c_void_p),
            t.ctypes.data_as(_ct.c_void_p), _ct.c_void_p(p_old_bar.data_ptr()),
            _ct.c_void_p(p_cur_bar.data_ptr()))
        if rc != 0:
            raise _lib.G3M3Error(f"g3m3_full_model_adjoint_inner returned {rc}")

    def corer(self, st, p_state):
        rc = self._lib.g3m3_full_model_corer(_ct.c_void_p(self.h), _ct.c_void_p(st.data_ptr()),
                                             _ct.c_void_p(p_state.data_ptr()))
        if rc != 0:
            raise _lib.G3M3Error(f"g3m3_full_model_corer returned {rc}")

    def corer_adjoint(self, s_bar, p_bar):
        rc = self._lib.g3m3_full_model_corer_adjoint(
            _ct.c_void_p(self.h), _ct.c_void_p(s_bar.data_ptr()), _ct.c_void_p(p_bar.data_ptr()))
        if rc != 0:
            raise _lib.G3M3Error(f"g3m3_full_model_corer_adjoint returned {rc}")

    def _ptrs(self, st):
        return (_ct.c_void_p * len(FULL_SAMPLED))(*[st[i].data_ptr() for i in FULL_SAMPLED])

    def sample(self, st, pos, out, gp):
        mn, mx, xs, re = self._grid
        _lib.call("sample_fields", pos.shape[0], pos, 1, mn, mx, xs, re, self.nx, self.ny,
                  self.nz, len(FULL_SAMPLED), self._ptrs(st), out, gp)

    def sample_adjoint(self, pos, s_bar, st_bar):
        mn, mx, xs, re = self._grid
        _lib.call("sample_fields_adjoint", pos.shape[0], pos, 1, mn, mx, xs, re, self.nx,
                  self.ny, self.nz, len(FULL_SAMPLED), self._ptrs(st_bar), s_bar)


def _full_function():
    import torch

    class FullRollout(torch.autograd.Function):
        """(ctl, state at step 0) -> the FULL_SAMPLED quantities at the craft
        after every ``every``-th of ``steps`` steps, shape (steps // every,
        npts, 8).  ctl is a float64 CPU tensor of 9 (rounded to float32 for
        the device); the state a (35, V) float32 device tensor.  Backward:
        the gradient in ctl (float64, accumulated in double on the device)
        and in the state, by checkpointed recompute as FieldRollout."""

        @staticmethod
        def forward(ctx, ctl, s_start, model, steps, every, pos, ckpt, keep):
            c = ctl.detach().cpu().numpy()
            ns = steps // every
            nq = len(FULL_SAMPLED)
            out = torch.empty((ns, pos.shape[0], nq), dtype=torch.float32, device="cuda")
            gp = torch.empty((pos.shape[0], 4), dtype=torch.int32, device="cuda")
            s = s_start.detach().clone()
            ns_ = torch.empty_like(s)
            ckpts = []
            last = (steps - 1) // ckpt * ckpt
            tapes = [model.new_tape() for _ in range(last, steps)] if keep else None
            for t in range(steps):
                if t % ckpt == 0:
                    ckpts.append(s.clone())
                model.step(s, ns_, c, tapes[t - last] if keep and t >= last else None)
                s, ns_ = ns_, s
                if (t + 1) % every == 0:
                    model.sample(s, pos, out[(t + 1) // every - 1], gp)
            ctx.model, ctx.steps, ctx.every, ctx.ckpt = model, steps, every, ckpt
            ctx.c, ctx.pos, ctx.ckpts, ctx.last_tapes = c, pos, ckpts, tapes
            ctx.ctl_meta = (ctl.dtype, ctl.device, ctl.shape)
            return out

        @staticmethod
        def backward(ctx, gs):
            model, steps, every, ckpt = ctx.model, ctx.steps, ctx.every, ctx.ckpt
            gs = gs.contiguous()
            s_bar = torch.zeros((NSTATE, model.V), dtype=torch.float32, device="cuda")
            tmp = torch.empty_like(s_bar)
            ctl_bar = torch.zeros(len(FULL_CTL_NAMES), dtype=torch.float64, device="cuda")
            for seg in reversed(range(0, steps, ckpt)):
                end = min(seg + ckpt, steps)
                tapes = ctx.last_tapes
                ctx.last_tapes = None
                if tapes is None:
                    # recompute the segment from its checkpoint (the last
                    # segment's tapes come from the forward pass)
                    tapes = [model.new_tape() for _ in range(seg, end)]
                    s = ctx.ckpts[seg // ckpt].clone()
                    ns_ = torch.empty_like(s)
                    for t in range(seg, end):
                        model.step(s, ns_, ctx.c, tapes[t - seg])
                        s, ns_ = ns_, s
                for t in reversed(range(seg, end)):
                    if (t + 1) % every == 0:
                        model.sample_adjoint(ctx.pos, gs[(t + 1) // every - 1], s_bar)
                    model.adjoint(tapes[t - seg], s_bar, tmp, ctl_bar, ctx.c)
                    s_bar, tmp = tmp, s_bar
                del tapes
            dtype, device, shape = ctx.ctl_meta
            g_ctl = ctl_bar.cpu().to(dtype).reshape(shape).to(device)
            return g_ctl, s_bar, None, None, None, None, None, None

    return FullRollout


_FULL = None


def full_rollout(ctl, s_start, model: FullModel, steps: int, every: int, pos, ckpt: int = 10):
    """Craft samples of the full model's rollout, differentiable in ctl and
    the initial state."""
    import torch
    global _FULL
    if _FULL is None:
        _FULL = _full_function()
    keep = torch.is_grad_enabled() and (ctl.requires_grad or s_start.requires_grad)
    return _FULL.apply(ctl, s_start, model, steps, every, pos, ckpt, keep)
