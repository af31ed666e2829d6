"""The single-subroutine adjoints of the ion push, the species Lapidus
smoothing and the species velocity (g3m3_push_ion_adjoint,
g3m3_lap_plasma_adjoint, g3m3_species_velocity_adjoint) against float64
vector-Jacobian products of the torch restatements (tests/ionref.py, pinned
bit for bit to the oracle in float32 by tests/test_ion_adjoint_host.py), on
the seeded inputs of the per-subroutine tests, with a random output
adjoint.  Bar: relative L2 error 1e-5 per adjoint array (float32 adjoint
against the exact float64 product of the same function)."""
import ctypes

import numpy as np
import pytest
import torch

from common import World
from ionref import FIELDS10, lap_plasma, push_ion, species_velocity

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _ptrs(ts):
    return np.array([t.data_ptr() for t in ts], np.uint64)


@pytest.mark.parametrize("sp,rm", [("q", "rmassq"), ("o", "rmasso")])
def test_push_ion_adjoint(sp, rm):
    from g3m3_amd import _lib
    w = World(20, 14, 11, 2, seed=83)
    w.s.update(box=2, delt=np.float32(0.037))
    b = 1
    nz, ny, nx = 11, 14, 20
    g = torch.Generator().manual_seed(5)
    src = [w.a["wrk" + sp + f][b] for f in FIELDS10]
    old = [w.a["old" + sp + f][b] for f in FIELDS10]
    out0 = [w.a[sp + f][b] for f in FIELDS10]
    fields3 = {k: w.a[k] for k in ("bsx", "bsy", "bsz", "btot", "efldx", "efldy", "efldz",
                                   "tvx", "tvy", "tvz")}
    obar = [torch.randn((nz, ny, nx), generator=g) for _ in range(10)]
    grd = [float(w.a[k][b]) for k in ("grd_xmin", "grd_xmax", "grd_ymin", "grd_ymax",
                                      "grd_zmin", "grd_zmax")]
    sc = dict(rmass=w.s[rm], delt=w.s["delt"], grav=w.s["grav"], re_equiv=w.s["re_equiv"],
              reynolds=w.s["reynolds"], grd=grd)
    # device
    D = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    dsrc, dold = [D(a) for a in src], [D(a) for a in old]
    d3 = {k: D(v) for k, v in fields3.items()}
    dob = [o.cuda() for o in obar]
    sbar = [torch.zeros((nz, ny, nx), device="cuda") for _ in range(10)]
    obarr = [torch.zeros((nz, ny, nx), device="cuda") for _ in range(10)]
    z3 = {k: torch.zeros((nz, ny, nx), device="cuda") for k in fields3}
    work = torch.zeros(15 * nx * ny * nz, device="cuda")
    _lib.call("push_ion_adjoint", _ptrs(dsrc), d3["bsx"], d3["bsy"], d3["bsz"], d3["btot"],
              d3["efldx"], d3["efldy"], d3["efldz"], d3["tvx"], d3["tvy"], d3["tvz"],
              _ptrs(dob), _ptrs(sbar), _ptrs(obarr), z3["bsx"], z3["bsy"], z3["bsz"],
              z3["btot"], z3["efldx"], z3["efldy"], z3["efldz"], z3["tvx"], z3["tvy"],
              z3["tvz"], nx, ny, nz, np.float32(sc["rmass"]), np.float32(sc["delt"]),
              np.float32(sc["grav"]), np.float32(sc["re_equiv"]), np.float32(sc["reynolds"]),
              *[np.float32(x) for x in grd], work)
    _lib.sync()
    # float64 reference
    T = lambda a: torch.from_numpy(np.asarray(a, np.float64)).requires_grad_(True)
    tsrc, told = [T(a) for a in src], [T(a) for a in old]
    t3 = {k: T(v) for k, v in fields3.items()}
    outs = push_ion([torch.from_numpy(np.asarray(a, np.float64)) for a in out0], told, tsrc,
                    [t3["bsx"], t3["bsy"], t3["bsz"]], t3["btot"],
                    [t3["efldx"], t3["efldy"], t3["efldz"]], [t3["tvx"], t3["tvy"], t3["tvz"]],
                    sc, torch.float64)
    L = sum((o * ob.double()).sum() for o, ob in zip(outs, obar))
    names = [f"src {f}" for f in FIELDS10] + [f"old {f}" for f in FIELDS10] + list(fields3)
    inputs = tsrc + told + list(t3.values())
    grads = torch.autograd.grad(L, inputs, allow_unused=True)
    got = [s.cpu().numpy() for s in sbar] + [o.cpu().numpy() for o in obarr] + \
        [z3[k].cpu().numpy() for k in fields3]
    errs = {}
    for nm, gr, gt in zip(names, grads, got):
        want = np.zeros(gt.shape) if gr is None else gr.numpy()
        if np.linalg.norm(want) == 0.0:
            assert np.abs(gt).max() == 0.0, nm
            continue
        errs[nm] = _rel(gt, want)
    print({k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    bad = {k: v for k, v in errs.items() if v > 1e-5}
    assert not bad, bad


def test_lap_plasma_and_velocity_adjoint():
    from g3m3_amd import _lib
    w = World(20, 14, 11, 2, seed=91)
    w.s.update(box=2, delt=np.float32(0.037))
    b = 1
    nz, ny, nx = 11, 14, 20
    g = torch.Generator().manual_seed(6)
    q = [w.a["h" + f][b] for f in FIELDS10]
    w0 = [w.a["wrkh" + f][b] for f in FIELDS10]
    wbar = [torch.randn((nz, ny, nx), generator=g) for _ in range(10)]
    chirho, chipxyz, delt, small = w.s["chirho"], w.s["chipxyz"], w.s["delt"], w.s["smallbit"]
    D = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    # v from the species (as the main program forms it before lap_plasma)
    dq = [D(a) for a in q]
    den = torch.clamp(dq[0], min=float(small))
    dv = [dq[4] / den, dq[5] / den, dq[6] / den]
    qbar = [torch.zeros((nz, ny, nx), device="cuda") for _ in range(10)]
    vbar = [torch.zeros((nz, ny, nx), device="cuda") for _ in range(3)]
    work = torch.zeros(6 * nx * ny * nz, device="cuda")
    dwb = [x.cuda() for x in wbar]
    _lib.call("lap_plasma_adjoint", _ptrs(dq), dv[0], dv[1], dv[2], _ptrs(dwb), _ptrs(qbar),
              vbar[0], vbar[1], vbar[2], nx, ny, nz, np.float32(chirho), np.float32(chipxyz),
              np.float32(delt), work)
    _lib.call("species_velocity_adjoint", dq[0], dq[4], dq[5], dq[6], vbar[0], vbar[1],
              vbar[2], qbar[0], qbar[4], qbar[5], qbar[6], nx * ny * nz, np.float32(small))
    _lib.sync()
    T = lambda a: torch.from_numpy(np.asarray(a, np.float64)).requires_grad_(True)
    tq = [T(a) for a in q]
    v = species_velocity(tq[0], [tq[4], tq[5], tq[6]], small, torch.float64)
    outs = lap_plasma(tq, [torch.from_numpy(np.asarray(a, np.float64)) for a in w0], v,
                      chirho, chipxyz, delt, torch.float64)
    L = sum((o * x.double()).sum() for o, x in zip(outs, wbar))
    grads = torch.autograd.grad(L, tq)
    errs = {f: _rel(qb.cpu().numpy(), gr.numpy()) for f, qb, gr in zip(FIELDS10, qbar, grads)}
    print({k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    bad = {k: v for k, v in errs.items() if v > 1e-5}
    assert not bad, bad
