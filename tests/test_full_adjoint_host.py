"""The torch restatement of the whole box-step (tests/fullref.py) against
the oracle, on the CPU.

* In float32 it is the oracle's box-step (OracleRun.box_step on a one-box
  run, every validation call included) bit for bit, over three steps of a
  20x16x12 box -- which pins the restatement, and shows that set_rho and
  set_speed_agrd change nothing on these states and that the step does not
  depend on the boundary cells the pushes leave in place (the oracle keeps
  the previous step's wrk arrays there, the restatement copies of the old
  state).
* Replaying the limiter branches a float32 step records gives the same
  bits (the float64 linearisation the device adjoint is compared with
  takes those branches).
* In float64 its gradient in each control matches central differences of
  itself (a check of the autograd path through the controls: the inflow
  values, the outer passes and eta).
"""
from __future__ import annotations

import numpy as np
import torch

from fullref import FullRef, oracle_run, oracle_step
from g3m3_amd.adjoint import FULL_CTL_NAMES, full_setup
from g3m3_amd.workload import FIELDS

SMALL = dict(nx=20, ny=16, nz=12, spacing=1.0)


def _bits(a):
    return np.ascontiguousarray(np.asarray(a, np.float32)).view(np.uint32)


def test_float32_restatement_equals_oracle_box_step():
    s = full_setup(**SMALL)
    r = oracle_run(s, s.ctl)
    ref = FullRef(s, torch.float32)
    st = ref.state(s.state)
    c = [torch.tensor(float(x), dtype=torch.float32) for x in s.ctl]
    names = FIELDS + ["wrkepres"]
    for step in range(3):
        want = oracle_step(r, s.scal["delt"])
        st = ref.step(st, c)
        bad = {names[i]: int(np.count_nonzero(_bits(st[i].numpy()) != _bits(want[i])))
               for i in range(35)}
        bad = {k: v for k, v in bad.items() if v}
        assert not bad, f"step {step + 1}: {bad}"
        # the step moves every field (the comparison is not of unchanged arrays)
        assert all(not np.array_equal(want[i], s.state[i]) for i in range(35))


def test_recorded_branches_replay_bit_for_bit():
    s = full_setup(**SMALL)
    ref = FullRef(s, torch.float32)
    c = [torch.tensor(float(x), dtype=torch.float32) for x in s.ctl]
    rec = []
    a = ref.step(ref.state(s.state), c, record=rec)
    b = ref.step(ref.state(s.state), c, codes=rec)
    assert len(rec) == 34 and all(len(r) == 3 for r in rec)
    for x, y in zip(a, b):
        assert np.array_equal(_bits(x.numpy()), _bits(y.numpy()))


def test_float64_control_gradient_matches_central_differences():
    s = full_setup(**SMALL)
    ref = FullRef(s, torch.float64)
    g = torch.Generator().manual_seed(21)
    w = [torch.randn((s.nz, s.ny, s.nx), generator=g, dtype=torch.float64) for _ in range(35)]
    c0 = torch.tensor(s.ctl.astype(np.float64))
    rec = []
    FullRef(s, torch.float32).step(FullRef(s, torch.float32).state(s.state),
                                    [torch.tensor(float(x)) for x in s.ctl], record=rec)

    def L(c):
        y = ref.step(ref.state(s.state), [c[i] for i in range(9)], codes=rec)
        return sum((a * b).sum() for a, b in zip(y, w))

    c = c0.clone().requires_grad_(True)
    (gc,) = torch.autograd.grad(L(c), [c])
    with torch.no_grad():
        for i in range(9):
            h = 1e-4 * max(abs(float(c0[i])), 0.1)
            cp, cm = c0.clone(), c0.clone()
            cp[i] += h
            cm[i] -= h
            fd = (float(L(cp)) - float(L(cm))) / (2 * h)
            assert abs(float(gc[i]) - fd) <= 1e-6 * max(abs(fd), 1e-3), \
                (FULL_CTL_NAMES[i], float(gc[i]), fd)
