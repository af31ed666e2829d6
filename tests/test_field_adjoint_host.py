"""CPU checks of the references the field-update adjoint is held to.

* The torch restatement (tests/fieldref.py) in float32 reproduces, bit for
  bit, the oracle's composition of the reference subroutines
  (tests/fieldoracle.py) -- so the float64 run of the same code is the
  reference model -- and its craft sampling matches the oracle's.
* Its float64 autograd gradient of a craft misfit matches float64 central
  differences, in every control and in a direction of the initial field
  (the difference step stays inside one branch of fcsmooth's limiter).
* The synthetic setup is deterministic and stays bounded over the steps
  the GPU tests take.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from g3m3_amd.adjoint import craft_positions, field_setup
from fieldoracle import FieldOracle
from fieldref import FieldRef

SMALL = dict(nx=17, ny=15, nz=11, spacing=1.0)


def _bits(a):
    return np.ascontiguousarray(np.asarray(a, np.float32)).view(np.uint32)


def test_setup_deterministic():
    a, b = field_setup(**SMALL), field_setup(**SMALL)
    for k in a.plasma:
        assert np.array_equal(_bits(a.plasma[k]), _bits(b.plasma[k]))
    assert np.array_equal(_bits(a.b_init), _bits(b.b_init))
    assert a.scal == b.scal


def test_ref_float32_equals_oracle_composition():
    s = field_setup(**SMALL)
    ora = FieldOracle(s)
    ref = FieldRef(s, torch.float32)
    ctl = np.array([2e-3, 0.05, -0.03, -0.24], np.float32)
    b_o = s.b_init.copy()
    b_r = ref.state(s.b_init)
    ct = torch.tensor(ctl, dtype=torch.float32)
    for step in range(3):
        b_o = ora.step(b_o, ctl)
        b_r = ref.step(b_r, ct)
        got = np.stack([x.numpy() for x in b_r])
        bad = int(np.count_nonzero(_bits(got) != _bits(b_o)))
        assert bad == 0, f"step {step + 1}: {bad} cells differ"


def test_ref_sampling_equals_oracle():
    s = field_setup(**SMALL)
    ora = FieldOracle(s)
    ref = FieldRef(s, torch.float32)
    pos = craft_positions(s, 16, seed=5)
    want, gp = ora.sample(s.b_init, pos)
    assert np.all(gp[:, 0] > 0)
    got = ref.sample(ref.state(s.b_init), ref.sample_plan(pos)).numpy()
    assert np.array_equal(_bits(got), _bits(want))


def _loss_fn(ref, s, pos, steps, every, target):
    plan = ref.sample_plan(pos)

    def loss(ctl, b):
        out = ref.rollout(ctl, b, steps, every, plan)
        return ((out - target) ** 2).sum()
    return loss, plan


def test_ref_float64_gradient_matches_central_differences():
    s = field_setup(**SMALL)
    ref = FieldRef(s, torch.float64)
    pos = craft_positions(s, 8, seed=6)
    steps, every = 6, 2
    ctl_true = torch.tensor([1e-3, 0.04, 0.02, -0.20], dtype=torch.float64)
    plan = ref.sample_plan(pos)
    with torch.no_grad():
        target = ref.rollout(ctl_true, ref.state(s.b_init), steps, every, plan)
    loss, _ = _loss_fn(ref, s, pos, steps, every, target)
    ctl0 = torch.tensor([2e-3, 0.0, 0.0, -0.24], dtype=torch.float64, requires_grad=True)
    b0 = tuple(x.clone().requires_grad_(True) for x in ref.state(s.b_init))
    L = loss(ctl0, b0)
    g_ctl, *g_b = torch.autograd.grad(L, [ctl0, *b0])
    rng = np.random.default_rng(7)
    v = tuple(torch.from_numpy(rng.standard_normal(x.shape)) for x in b0)
    with torch.no_grad():
        for i in range(4):
            h = 1e-4 * max(abs(float(ctl0[i])), 1e-2)
            cp, cm = ctl0.detach().clone(), ctl0.detach().clone()
            cp[i] += h
            cm[i] -= h
            fd = (float(loss(cp, ref.state(s.b_init))) - float(loss(cm, ref.state(s.b_init)))) / (2 * h)
            assert abs(fd - float(g_ctl[i])) <= 1e-5 * max(abs(fd), 1e-8) + 1e-12, (i, fd, g_ctl[i])
        # fcsmooth's limiter makes the misfit piecewise smooth; a random
        # direction in every cell meets a branch switch within ~1e-6, so the
        # difference step is small (float64 keeps it exact enough)
        h = 1e-7
        bp = tuple(x.detach() + h * d for x, d in zip(b0, v))
        bm = tuple(x.detach() - h * d for x, d in zip(b0, v))
        fd = (float(loss(ctl0.detach(), bp)) - float(loss(ctl0.detach(), bm))) / (2 * h)
        ad = sum(float((g * d).sum()) for g, d in zip(g_b, v))
        assert abs(fd - ad) <= 1e-5 * abs(fd), (fd, ad)


def test_setup_bounded_over_steps():
    s = field_setup(nx=41, ny=41, nz=21, spacing=0.5)
    ref = FieldRef(s, torch.float32)
    b = ref.state(s.b_init)
    ct = torch.tensor(s.ctl, dtype=torch.float32)
    m0 = max(float(x.abs().max()) for x in b)
    for _ in range(20):
        b = ref.step(b, ct)
    m = max(float(x.abs().max()) for x in b)
    assert np.isfinite(m) and m < 3.0 * m0


def test_recorded_branches_reproduce_the_step():
    """A step that takes the branches a float32 step recorded is that step
    (float32, bit for bit), and in float64 it is differentiable with the
    same gradient as the float64 step wherever no branch differs."""
    s = field_setup(**SMALL)
    r32 = FieldRef(s, torch.float32)
    ct = torch.tensor(s.ctl, dtype=torch.float32)
    b = r32.state(s.b_init)
    rec = []
    y1 = r32.step(b, ct, record=rec)
    y2 = r32.step(b, ct, codes=rec)
    for a, c in zip(y1, y2):
        assert np.array_equal(_bits(a.numpy()), _bits(c.numpy()))
    assert len(rec) == 3 and all(len(r) == 3 for r in rec)
    assert any(int((r != 0).sum()) > 0 for rr in rec for r in rr)
