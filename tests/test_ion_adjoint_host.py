"""CPU pins of the torch restatements the ion-push and Lapidus adjoints are
held to (tests/ionref.py): in float32 they equal the oracle's push_ion and
lap_plasma bit for bit on the same inputs, and the species velocity is
p / max(rho, smallbit)."""
import numpy as np
import pytest
import torch

from common import World
from ionref import FIELDS10, lap_plasma, push_ion
from oracle import restate as O
from oracle.argnames import lap_plasma_names, push_ion_names


def _bits(a):
    return np.ascontiguousarray(np.asarray(a, np.float32)).view(np.uint32)


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a))


@pytest.mark.parametrize("dims", [(13, 11, 9), (20, 14, 11)])
@pytest.mark.parametrize("sp,rm", [("q", "rmassq"), ("o", "rmasso")])
def test_push_ion_float32_equals_oracle(dims, sp, rm):
    w = World(*dims, 2, seed=41 + dims[0])
    w.s.update(box=2, delt=np.float32(0.037))
    b = 1
    src = [_t(w.a["wrk" + sp + f][b]) for f in FIELDS10]
    old = [_t(w.a["old" + sp + f][b]) for f in FIELDS10]
    out0 = [_t(w.a[sp + f][b]) for f in FIELDS10]
    bs = [_t(w.a[k]) for k in ("bsx", "bsy", "bsz")]
    ef = [_t(w.a[k]) for k in ("efldx", "efldy", "efldz")]
    tv = [_t(w.a[k]) for k in ("tvx", "tvy", "tvz")]
    grd = [float(w.a[g][b]) for g in ("grd_xmin", "grd_xmax", "grd_ymin", "grd_ymax",
                                      "grd_zmin", "grd_zmax")]
    sc = dict(rmass=w.s[rm], delt=w.s["delt"], grav=w.s["grav"], re_equiv=w.s["re_equiv"],
              reynolds=w.s["reynolds"], grd=grd)
    got = push_ion(out0, old, src, bs, _t(w.a["btot"]), ef, tv, sc, torch.float32)
    wo = w.copy()
    O.push_ion(*wo.args(push_ion_names("", "old", "wrk", sp, rm)))
    for f, g in zip(FIELDS10, got):
        assert np.array_equal(_bits(g.numpy()), _bits(wo.a[sp + f][b])), f


def test_lap_plasma_float32_equals_oracle():
    w = World(13, 11, 9, 2, seed=7)
    w.s.update(box=2, delt=np.float32(0.037))
    b = 1
    q = [_t(w.a["h" + f][b]) for f in FIELDS10]
    w0 = [_t(w.a["wrkh" + f][b]) for f in FIELDS10]
    v = [_t(w.a[k]) for k in ("vvx", "vvy", "vvz")]
    got = lap_plasma(q, w0, v, w.s["chirho"], w.s["chipxyz"], w.s["delt"], torch.float32)
    wo = w.copy()
    O.lap_plasma(*wo.args(lap_plasma_names("h", "", "wrk")))
    for f, g in zip(FIELDS10, got):
        assert np.array_equal(_bits(g.numpy()), _bits(wo.a["wrkh" + f][b])), f
