"""The field-update adjoint on the device (g3m3_field_model_*,
g3m3_sample_fields_adjoint, g3m3_amd.adjoint.rollout).

* The device step is the oracle's composition of the reference subroutines
  bit for bit (tests/fieldoracle.py).
* The sampling adjoint passes the dot-product test against the sampling.
* One step's adjoint equals the float64 vector-Jacobian product of the
  torch restatement (tests/fieldref.py) at the same state, taking the
  limiter branches of the float32 step.
* The gradient of a craft misfit over a checkpointed rollout, in the four
  controls, matches the float64 autograd gradient and float64 central
  differences of the restatement within 1e-3; in the controls and the
  initial field it matches float64 vector-Jacobian products chained along
  the device's own states (with the float32 limiter branches) within 1e-5;
  it does not depend on the checkpoint interval.
"""
from __future__ import annotations

import sys
import time

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from g3m3_amd import _lib
from g3m3_amd.adjoint import FieldModel, craft_positions, field_setup, rollout
from fieldoracle import FieldOracle
from fieldref import FieldRef

SMALL = dict(nx=17, ny=15, nz=11, spacing=1.0)
MID = dict(nx=41, ny=41, nz=21, spacing=0.5)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _bits(a):
    return np.ascontiguousarray(np.asarray(a, np.float32)).view(np.uint32)


@pytest.mark.parametrize("dims", [SMALL, MID], ids=["17x15x11", "41x41x21"])
def test_step_equals_oracle(dims):
    _say("field step vs oracle", dims)
    s = field_setup(**dims)
    m = FieldModel(s)
    ora = FieldOracle(s)
    ctl = np.array([2e-3, 0.05, -0.03, -0.24], np.float32)
    b = m.new_state()
    nb = torch.empty_like(b)
    want = s.b_init.copy()
    for step in range(3):
        m.step(b, nb, ctl)
        b, nb = nb, b
        want = ora.step(want, ctl)
        got = b.cpu().numpy().reshape(want.shape)
        bad = int(np.count_nonzero(_bits(got) != _bits(want)))
        assert bad == 0, f"step {step + 1}: {bad} cells differ"
    m.close()


def test_step_with_tape_equals_without():
    s = field_setup(**SMALL)
    m = FieldModel(s)
    b = m.new_state()
    o1, o2 = torch.empty_like(b), torch.empty_like(b)
    m.step(b, o1, s.ctl)
    m.step(b, o2, s.ctl, m.new_tape())
    torch.cuda.synchronize()
    assert torch.equal(o1.view(torch.int32), o2.view(torch.int32))
    m.close()


def test_sample_adjoint_dot_product():
    _say("sample adjoint dot product")
    s = field_setup(**MID)
    m = FieldModel(s)
    pos = torch.from_numpy(craft_positions(s, 32, seed=11)).cuda()
    g = torch.Generator().manual_seed(3)
    b = torch.randn((3, m.V), generator=g).cuda()
    w = torch.randn((32, 3), generator=g).cuda()
    out = torch.empty((32, 3), device="cuda")
    gp = torch.empty((32, 4), dtype=torch.int32, device="cuda")
    m.sample(b, pos, out, gp)
    bb = torch.zeros_like(b)
    m.sample_adjoint(pos, w, bb)
    _lib.sync()
    assert bool((gp[:, 0] > 0).all())
    lhs = float((out.double() * w.double()).sum())
    rhs = float((b.double() * bb.double()).sum())
    assert abs(lhs - rhs) <= 1e-5 * max(abs(lhs), 1.0), (lhs, rhs)
    m.close()


def _branches(s, b_h, ctl):
    """fcsmooth's limiter branches of the float32 step from b_h (the device
    step's own, which the float32 restatement reproduces bit for bit)."""
    r32 = FieldRef(s, torch.float32)
    rec = []
    r32.step(r32.state(b_h), torch.tensor(np.asarray(ctl, np.float32)), record=rec)
    return rec


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_step_adjoint_equals_float64_vjp():
    _say("step adjoint vs float64 vjp")
    s = field_setup(**MID)
    m = FieldModel(s)
    ctl = np.array([2e-3, 0.05, -0.03, -0.24], np.float32)
    b = m.new_state()
    nb = torch.empty_like(b)
    for _ in range(2):
        m.step(b, nb, ctl)
        b, nb = nb, b
    tape = m.new_tape()
    m.step(b, nb, ctl, tape)
    g = torch.Generator().manual_seed(4)
    w = torch.randn((3, m.V), generator=g)
    wd = w.cuda()
    bb = torch.empty_like(b)
    cb = torch.zeros(4, dtype=torch.float64, device="cuda")
    m.adjoint(tape, wd, bb, cb, ctl)
    torch.cuda.synchronize()
    ref = FieldRef(s, torch.float64)
    b_h = b.cpu().numpy().reshape(3, s.nz, s.ny, s.nx)
    codes = _branches(s, b_h, ctl)
    x = tuple(t.requires_grad_(True) for t in ref.state(b_h))
    c64 = torch.tensor(ctl.astype(np.float64), requires_grad=True)
    y = ref.step(x, c64, codes=codes)
    wr = w.double().reshape(3, s.nz, s.ny, s.nx)
    L = sum((yy * ww).sum() for yy, ww in zip(y, wr))
    gc, *gx = torch.autograd.grad(L, [c64, *x])
    got_b = bb.cpu().numpy().reshape(3, -1)
    want_b = np.stack([t.numpy().reshape(-1) for t in gx])
    eb = _rel(got_b, want_b)
    ec = [abs(float(cb[i]) - float(gc[i])) / max(abs(float(gc[i])), 1e-12) for i in range(4)]
    _say("rel err b:", eb, "ctl:", ec)
    assert eb < 1e-5
    assert max(ec) < 1e-5
    m.close()


def _misfit_case(dims=MID, steps=20, every=5, ncraft=8):
    s = field_setup(**dims)
    m = FieldModel(s)
    pos_h = craft_positions(s, ncraft, seed=12)
    pos = torch.from_numpy(pos_h).cuda()
    ctl_true = torch.tensor([1e-3, 0.04, 0.02, -0.20], dtype=torch.float64)
    with torch.no_grad():
        target = rollout(ctl_true, m.new_state(), m, steps, every, pos).clone()
    # observation noise of 5% of the samples' rms: residuals well above the
    # float32 rounding of the samples themselves, so the float32 misfit's
    # gradient can be held to 1e-3
    g = torch.Generator().manual_seed(13)
    rms = float(target.pow(2).mean().sqrt())
    target += (0.05 * rms) * torch.randn(target.shape, generator=g).cuda()
    return s, m, pos_h, pos, target


def _device_grad(m, pos, target, ctl0, steps, every, ckpt):
    ctl = ctl0.clone().requires_grad_(True)
    b0 = m.new_state().requires_grad_(True)
    out = rollout(ctl, b0, m, steps, every, pos, ckpt)
    L = ((out - target) ** 2).sum()
    L.backward()
    torch.cuda.synchronize()
    return float(L.detach()), ctl.grad.numpy().copy(), b0.grad.cpu().numpy().copy()


def test_rollout_gradient_matches_float64():
    _say("rollout gradient vs float64 reference and central differences")
    steps, every = 20, 5
    s, m, pos_h, pos, target = _misfit_case(steps=steps, every=every)
    ctl0 = torch.tensor([2e-3, 0.0, 0.0, -0.24], dtype=torch.float64)
    L, g_ctl, g_b = _device_grad(m, pos, target, ctl0, steps, every, 10)
    ref = FieldRef(s, torch.float64)
    plan = ref.sample_plan(pos_h)
    tgt = target.cpu().double()

    def loss(c, b):
        return ((ref.rollout(c, b, steps, every, plan) - tgt) ** 2).sum()

    c64 = ctl0.clone().requires_grad_(True)
    x = tuple(t.requires_grad_(True) for t in ref.state(s.b_init))
    L64 = loss(c64, x)
    gc, *gx = torch.autograd.grad(L64, [c64, *x])
    gx = np.stack([t.numpy().reshape(-1) for t in gx])
    _say("loss", L, float(L64.detach()))
    err_c = [abs(g_ctl[i] - float(gc[i])) / abs(float(gc[i])) for i in range(4)]
    err_b = _rel(g_b.reshape(3, -1), gx)
    _say("vs float64 autograd: ctl", err_c, "b", err_b)
    # The float64 model is a different function at fcsmooth's limiter
    # kinks: where a limiter argument rounds to an exact tie in float32 (a
    # difference of equal float32 values) but not in float64, the two take
    # different branches and the gradient in b differs in a few cells
    # (measured: ten cells around one craft hold 86% of the difference).
    # The controls, which sum over the whole box, stay within 1e-3; b is
    # held to 2e-3 here, and below to 1e-5 against float64 vector-Jacobian
    # products along the device's own states with the float32 branches.
    assert max(err_c) < 1e-3 and err_b < 2e-3
    # float64 vector-Jacobian products chained along the device's trajectory
    st = [m.new_state()]
    nb = torch.empty_like(st[0])
    c32 = ctl0.numpy().astype(np.float32)
    samples = []
    gpd = torch.empty((pos.shape[0], 4), dtype=torch.int32, device="cuda")
    for t in range(steps):
        m.step(st[-1], nb, c32)
        st.append(nb.clone())
        if (t + 1) % every == 0:
            o = torch.empty((pos.shape[0], 3), device="cuda")
            m.sample(st[-1], pos, o, gpd)
            samples.append(o)
    res = 2.0 * (torch.stack(samples).cpu().double() - tgt)
    shp = (3, s.nz, s.ny, s.nx)
    lam = tuple(torch.zeros(shp[1:], dtype=torch.float64) for _ in range(3))
    gc_chain = torch.zeros(4, dtype=torch.float64)
    cc = torch.tensor(c32.astype(np.float64), requires_grad=True)
    for t in reversed(range(steps)):
        if (t + 1) % every == 0:
            xs = tuple(a.requires_grad_(True) for a in ref.state(st[t + 1].cpu().numpy().reshape(shp)))
            sm = ref.sample(xs, plan)
            gs = torch.autograd.grad((sm * res[(t + 1) // every - 1]).sum(), xs)
            lam = tuple(a + b for a, b in zip(lam, gs))
        b_h = st[t].cpu().numpy().reshape(shp)
        xs = tuple(a.requires_grad_(True) for a in ref.state(b_h))
        y = ref.step(xs, cc, codes=_branches(s, b_h, c32))
        g1, *gl = torch.autograd.grad(sum((a * b).sum() for a, b in zip(y, lam)), [cc, *xs])
        gc_chain += g1
        lam = tuple(gl)
    lam = np.stack([a.numpy().reshape(-1) for a in lam])
    err_chain = _rel(g_b.reshape(3, -1), lam)
    err_cc = [abs(g_ctl[i] - float(gc_chain[i])) / abs(float(gc_chain[i])) for i in range(4)]
    _say("vs float64 adjoint along the device states: ctl", err_cc, "b", err_chain)
    assert err_chain < 1e-5 and max(err_cc) < 1e-5
    # float64 central differences of the restatement in each control.  The
    # step stays inside one branch of every limiter: at 1e-5 (1e-3 of the
    # scale) the sbx difference already crosses kinks and is 2.5% off, at
    # 1e-8 the float64 rounding of the misfit is ~3e-5 of these gradients.
    with torch.no_grad():
        for i in range(4):
            h = 1e-6 * max(abs(float(ctl0[i])), 1e-2)
            cp, cm = ctl0.clone(), ctl0.clone()
            cp[i] += h
            cm[i] -= h
            fd = (float(loss(cp, ref.state(s.b_init))) - float(loss(cm, ref.state(s.b_init)))) / (2 * h)
            _say("control", i, "device", g_ctl[i], "fd", fd)
            assert abs(g_ctl[i] - fd) <= 1e-3 * abs(fd), (i, g_ctl[i], fd)
    m.close()


def test_gradient_independent_of_checkpoint_interval():
    _say("checkpoint intervals")
    steps, every = 20, 5
    s, m, pos_h, pos, target = _misfit_case(steps=steps, every=every)
    ctl0 = torch.tensor([2e-3, 0.0, 0.0, -0.24], dtype=torch.float64)
    res = [_device_grad(m, pos, target, ctl0, steps, every, k) for k in (1, 7, 10, 20)]
    L0, c0, b0 = res[0]
    for L, c, b in res[1:]:
        assert L == L0
        assert np.allclose(c, c0, rtol=1e-5, atol=0)
        assert _rel(b, b0) < 1e-5
    m.close()


def test_adjoint_cost_ratio():
    _say("adjoint cost")
    steps, every = 20, 5
    s, m, pos_h, pos, target = _misfit_case(steps=steps, every=every)
    ctl0 = torch.tensor([2e-3, 0.0, 0.0, -0.24], dtype=torch.float64)
    _device_grad(m, pos, target, ctl0, steps, every, 10)     # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with torch.no_grad():
        rollout(ctl0, m.new_state(), m, steps, every, pos)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    _device_grad(m, pos, target, ctl0, steps, every, 10)
    t2 = time.perf_counter()
    fwd, both = t1 - t0, t2 - t1
    _say(f"forward {fwd * 1e3:.2f} ms, forward+backward {both * 1e3:.2f} ms, "
         f"backward/forward {(both - fwd) / fwd:.2f}")
    assert np.isfinite(both)
    m.close()
