"""The nested-box restatement (tests/fullref.py NestedRef: FullRef's
box-step with FullRef.flanks and FullRef.corer on
g3m3_amd.adjoint_nested.nested_schedule) against the oracle, on the CPU.

* The schedule's box-steps are the oracle's full step's (OracleRun.full_step,
  code/3d_main.f90:1925-3095), box and times, for two and three boxes.
* In float32 the restatement is the oracle's full step bit for bit, every
  array of every box, over two full steps of two and of three nested boxes
  -- which pins FullRef.flanks (code/sub_flanks.f90), FullRef.corer
  (code/sub_corer.f90), the flanks' time weights and the order of the
  box-steps and corer calls, and shows that the validation calls change
  nothing on these states.
* In float64 the gradient of a sum over the boxes' states after one full
  step matches central differences of itself in eta and in the upstream
  density (through the outer box's inflow, the flanks and the corer).
* The schedule's shape for four boxes (visits, corer calls, time weights),
  its refusal of other time-step ratios, and the interpolation that seeds
  the inner boxes (against scipy's first-order map_coordinates).
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from fullref import NestedRef, nested_oracle_full_step, nested_oracle_run
from g3m3_amd.adjoint_nested import nested_schedule, nested_setup

DIMS = dict(nx=14, ny=12, nz=10, spacing=1.0)


def _bits(a):
    return np.ascontiguousarray(np.asarray(a, np.float32)).view(np.uint32)


@pytest.mark.parametrize("nbox", [2, 3])
def test_schedule_equals_oracle_full_step(nbox):
    ss = nested_setup(nbox, **DIMS)
    ts = [s.scal["delt"] for s in ss]
    ev = nested_schedule(ts)
    r = nested_oracle_run(ss, ss[-1].ctl)
    nested_oracle_full_step(r, ss)
    t_new = [0.0] * nbox
    mine = []
    for e in ev:
        if e[0] == "step":
            b = e[1]
            mine.append((b + 1, t_new[b], t_new[b] + float(np.float32(ts[b]))))
            t_new[b] = mine[-1][2]
    assert mine == [(b, float(a), float(c)) for b, a, c in r.step_log]
    assert [e for e in ev if e[0] == "corer"][-(nbox - 1):] == [("corer", b)
                                                               for b in range(nbox - 1)]


@pytest.mark.parametrize("nbox", [2, 3])
def test_float32_nested_restatement_equals_oracle(nbox):
    ss = nested_setup(nbox, **DIMS)
    r = nested_oracle_run(ss, ss[-1].ctl)
    ref = NestedRef(ss, torch.float32)
    cur = ref.states()
    c = [torch.tensor(float(x), dtype=torch.float32) for x in ss[-1].ctl]
    for step in range(2):
        want = nested_oracle_full_step(r, ss)
        with torch.no_grad():
            cur = ref.full_step(cur, c)
        for b in range(nbox):
            got = np.stack([x.numpy() for x in cur[b]])
            bad = {i: int(np.count_nonzero(_bits(got[i]) != _bits(want[b][i])))
                   for i in range(got.shape[0])}
            bad = {k: v for k, v in bad.items() if v}
            assert not bad, f"full step {step + 1}, box {b + 1}: differing cells {bad}"


def test_float64_nested_gradient_matches_central_differences():
    ss = nested_setup(2, **DIMS)
    ref = NestedRef(ss, torch.float64)
    g = torch.Generator().manual_seed(3)
    w = [[torch.randn(x.shape, generator=g, dtype=torch.float64) for x in st]
         for st in ref.states()]

    def loss(c):
        out = ref.full_step(ref.states(), [c[i] for i in range(9)])
        return sum((x * y).sum() for st, ws in zip(out, w) for x, y in zip(st, ws))

    c0 = torch.tensor(ss[-1].ctl.astype(np.float64))
    c = c0.clone().requires_grad_(True)
    (gr,) = torch.autograd.grad(loss(c), [c])
    with torch.no_grad():
        for i in (0, 4):
            h = 1e-6 * max(abs(float(c0[i])), 0.1)
            cp, cm = c0.clone(), c0.clone()
            cp[i] += h
            cm[i] -= h
            fd = (float(loss(cp)) - float(loss(cm))) / (2 * h)
            assert abs(float(gr[i]) - fd) <= 1e-5 * abs(fd), (i, float(gr[i]), fd)


def test_schedule_shape_four_boxes():
    """Four boxes, each time step twice the inner one: 8 visits per full
    step; box b+1 steps 2^(3-b) times, every box but the outermost catches
    up with its parent after each of the parent's steps (1 + 2 + 4 corer
    calls), then every inner box is restricted once more."""
    ev = nested_schedule([0.125, 0.25, 0.5, 1.0])
    steps = [e[1] for e in ev if e[0] == "step"]
    assert [steps.count(b) for b in range(4)] == [8, 4, 2, 1]
    corers = [e[1] for e in ev if e[0] == "corer"]
    assert [corers.count(b) for b in range(3)] == [4 + 1, 2 + 1, 1 + 1]
    assert ev[0] == ("step", 3, None)
    # the flanks' time weights stay in [0, 1] and sum to 1
    for e in ev:
        if e[0] == "step" and e[2] is not None:
            t1a, t2a, t1b, t2b = (float(x) for x in e[2])
            assert 0.0 <= t1a <= 1.0 and 0.0 <= t1b <= 1.0
            assert t1a + t2a == 1.0 and t1b + t2b == 1.0


def test_schedule_refuses_other_ratios():
    with pytest.raises(ValueError):
        nested_schedule([0.1, 0.3])


def test_inner_box_interpolation_matches_scipy():
    """nested_setup's trilinear interpolation of the parent state (before
    its smooth variations) is scipy's first-order map_coordinates."""
    from scipy.ndimage import map_coordinates
    from g3m3_amd.adjoint_nested import _trilinear
    rng = np.random.default_rng(9)
    src = rng.normal(size=(2, 9, 8, 7))
    x, y, z = np.linspace(0.5, 5.5, 6), np.linspace(1.0, 6.0, 5), np.linspace(0.0, 8.0, 4)
    got = _trilinear(src, x, y, z)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    for a in range(2):
        want = map_coordinates(src[a], [Z, Y, X], order=1)
        assert np.allclose(got[a], want, rtol=0, atol=1e-12)
