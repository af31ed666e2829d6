"""The whole-box-step adjoint on the device (g3m3_full_model_*,
g3m3_amd.adjoint.full_rollout).

* The device step is the oracle's box-step (OracleRun.box_step, one box,
  mbndry = 0, fixed time step) bit for bit, all 35 state arrays.
* One step's adjoint equals the float64 vector-Jacobian product of the
  torch restatement (tests/fullref.py, pinned to the oracle in float32 by
  tests/test_full_adjoint_host.py) at the same state, linearised at the
  float32 step's values (fullref.Float32Values: the float32 intermediates
  and limiter branches the device differentiates): relative error 1e-5
  per state array group and per control.
* The gradient of a craft misfit (density, momentum, electron pressure and
  field sampled at virtual spacecraft) over a checkpointed rollout matches
  float64 vector-Jacobian products chained along the device's own states
  (linearised at its float32 values) within 1e-5 or 3x the error of torch's
  float32 autograd chained the same way, and does not depend on the
  checkpoint interval.  Against float64 central differences of the
  restatement it is within 1e-3 per control plus the difference between
  the float64 gradients linearised at float32 and at float64 values (the
  float32 and float64 models differ; measured up to 3e-3 of the gradient
  for the controls whose gradient is a cancelling sum); that difference is
  held below 5e-3, and the float64 gradient within 2e-4 of the differences.
"""
from __future__ import annotations

import sys
import time

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from g3m3_amd.adjoint import (FULL_CTL_NAMES, FULL_SAMPLED, NSTATE, FullModel,
                              craft_positions, full_rollout, full_setup)
from fullref import Float32Values, FullRef, oracle_run, oracle_step

SMALL = dict(nx=20, ny=16, nz=12, spacing=1.0)
MID = dict(nx=41, ny=41, nz=21, spacing=0.5)
GROUPS = {"q": range(0, 10), "h": range(10, 20), "o": range(20, 30), "epres": [30],
          "b": range(31, 34), "wrkepres": [34]}


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _bits(a):
    return np.ascontiguousarray(np.asarray(a, np.float32)).view(np.uint32)


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _ctl_list(c, dtype):
    return [c[i] for i in range(len(FULL_CTL_NAMES))] if torch.is_tensor(c) else \
        [torch.tensor(float(x), dtype=dtype) for x in c]


@pytest.mark.parametrize("dims", [SMALL, MID], ids=["20x16x12", "41x41x21"])
def test_full_step_equals_oracle(dims):
    _say("full step vs oracle", dims)
    s = full_setup(**dims)
    m = FullModel(s)
    r = oracle_run(s, s.ctl)
    st = m.new_state()
    nst = torch.empty_like(st)
    for step in range(3):
        m.step(st, nst, s.ctl)
        st, nst = nst, st
        want = oracle_step(r, s.scal["delt"])
        got = st.cpu().numpy().reshape(want.shape)
        bad = {i: int(np.count_nonzero(_bits(got[i]) != _bits(want[i]))) for i in range(NSTATE)}
        bad = {k: v for k, v in bad.items() if v}
        assert not bad, f"step {step + 1}: arrays with differing cells {bad}"
    m.close()


def test_full_step_with_tape_equals_without():
    s = full_setup(**SMALL)
    m = FullModel(s)
    st = m.new_state()
    o1, o2 = torch.empty_like(st), torch.empty_like(st)
    m.step(st, o1, s.ctl)
    m.step(st, o2, s.ctl, m.new_tape())
    torch.cuda.synchronize()
    assert torch.equal(o1.view(torch.int32), o2.view(torch.int32))
    m.close()


def _branches(s, st_h, ctl):
    """The limiter branches of the float32 step from st_h (the device's
    own step, which the float32 restatement reproduces bit for bit)."""
    r32 = FullRef(s, torch.float32)
    rec = []
    r32.step(r32.state(st_h), _ctl_list(np.asarray(ctl, np.float32), torch.float32), record=rec)
    return rec


@pytest.mark.parametrize("dims", [SMALL, MID], ids=["20x16x12", "41x41x21"])
def test_full_step_adjoint_equals_float64_vjp(dims):
    _say("full step adjoint vs float64 vjp", dims)
    s = full_setup(**dims)
    m = FullModel(s)
    ctl = s.ctl.copy()
    st = m.new_state()
    nst = torch.empty_like(st)
    for _ in range(2):
        m.step(st, nst, ctl)
        st, nst = nst, st
    tape = m.new_tape()
    m.step(st, nst, ctl, tape)
    g = torch.Generator().manual_seed(8)
    w = torch.randn((NSTATE, m.V), generator=g)
    sb = torch.empty_like(st)
    cb = torch.zeros(len(FULL_CTL_NAMES), dtype=torch.float64, device="cuda")
    m.adjoint(tape, w.cuda(), sb, cb, ctl)
    torch.cuda.synchronize()
    shp = (NSTATE, s.nz, s.ny, s.nx)
    st_h = st.cpu().numpy().reshape(shp)
    ref = FullRef(s, torch.float64)
    x = [t.requires_grad_(True) for t in ref.state(st_h)]
    c64 = torch.tensor(ctl.astype(np.float64), requires_grad=True)
    codes = _branches(s, st_h, ctl)
    with Float32Values():
        y = ref.step(x, _ctl_list(c64, torch.float64), codes=codes)
    wr = w.double().reshape(shp)
    L = sum((yy * ww).sum() for yy, ww in zip(y, wr))
    gc, *gx = torch.autograd.grad(L, [c64, *x], allow_unused=True)
    got = sb.cpu().numpy().reshape(NSTATE, -1)
    want = np.stack([np.zeros(s.nz * s.ny * s.nx) if t is None else t.numpy().reshape(-1)
                     for t in gx])
    # torch's float32 autograd of the same restatement: the accuracy a
    # float32 reverse pass of this step reaches
    r32 = FullRef(s, torch.float32)
    x32 = [t.requires_grad_(True) for t in r32.state(st_h)]
    c32 = torch.tensor(ctl, requires_grad=True)
    y32 = r32.step(x32, _ctl_list(c32, torch.float32), codes=codes)
    L32 = sum((yy * ww).sum() for yy, ww in zip(y32, w.reshape(shp)))
    gc32, *gx32 = torch.autograd.grad(L32, [c32, *x32], allow_unused=True)
    t32 = np.stack([np.zeros(s.nz * s.ny * s.nx) if t is None else t.double().numpy().reshape(-1)
                    for t in gx32])
    errs = {k: _rel(got[list(v)], want[list(v)]) for k, v in GROUPS.items()}
    errs32 = {k: _rel(t32[list(v)], want[list(v)]) for k, v in GROUPS.items()}
    ec = {FULL_CTL_NAMES[i]: abs(float(cb[i]) - float(gc[i])) / max(abs(float(gc[i])), 1e-12)
          for i in range(len(FULL_CTL_NAMES))}
    _say("rel err by group:", {k: f"{v:.1e}" for k, v in errs.items()})
    _say("torch float32 autograd:", {k: f"{v:.1e}" for k, v in errs32.items()})
    _say("ctl:", {k: f"{v:.1e}" for k, v in ec.items()})
    # Bar: 1e-5, or 3x what float32 autograd of the restatement reaches.  The
    # field's adjoint needs the latter: in push_ion's momentum update the
    # E-field term and the ion v x B term nearly cancel (E ~ -ve x B), and
    # the two large, opposite contributions they send to b's adjoint leave
    # a float32 rounding residual of ~1e-4 of the net (torch's float32
    # autograd shows the same).
    bad = {k: v for k, v in errs.items() if v > max(1e-5, 3.0 * errs32[k])}
    assert not bad, (bad, errs32)
    assert max(ec.values()) < 1e-5, ec
    m.close()


def _misfit_case(dims=SMALL, steps=8, every=4, ncraft=8):
    s = full_setup(**dims)
    m = FullModel(s)
    pos_h = craft_positions(s, ncraft, seed=14)
    # craft 1.5 to 5 cells downstream of the inflow face, where the upstream
    # values reach them within the rollout
    h = float(s.grid["xspac"][0])
    x0 = float(s.grid["grid_minvals"][0][0])
    rng = np.random.default_rng(5)
    pos_h[:, 0] = (x0 + h * rng.uniform(1.5, 5.0, size=ncraft)).astype(np.float32)
    pos = torch.from_numpy(pos_h).cuda()
    ctl_true = torch.tensor(s.ctl.astype(np.float64))
    ctl_true[0] = 5e-4
    ctl_true[3] = -0.20
    ctl_true[4] *= 1.05
    ctl_true[5] *= 0.97
    with torch.no_grad():
        target = full_rollout(ctl_true, m.new_state(), m, steps, every, pos).clone()
    # 5% observation noise per sampled quantity: residuals well above the
    # float32 rounding of the samples
    g = torch.Generator().manual_seed(15)
    rms = target.pow(2).mean(dim=(0, 1)).sqrt()
    target += 0.05 * rms * torch.randn(target.shape, generator=g).cuda()
    # the quantities differ in scale by orders of magnitude: weight each by
    # its inverse rms
    wt = (1.0 / rms).cuda()
    return s, m, pos_h, pos, target, wt


def _device_grad(m, pos, target, wt, ctl0, steps, every, ckpt):
    ctl = ctl0.clone().requires_grad_(True)
    s0 = m.new_state().requires_grad_(True)
    out = full_rollout(ctl, s0, m, steps, every, pos, ckpt)
    L = (((out - target) * wt) ** 2).sum()
    L.backward()
    torch.cuda.synchronize()
    return float(L.detach()), ctl.grad.numpy().copy(), s0.grad.cpu().numpy().copy()


def test_full_rollout_gradient():
    _say("full rollout gradient: chained float64 vjps and central differences")
    steps, every = 8, 4
    s, m, pos_h, pos, target, wt = _misfit_case(steps=steps, every=every)
    ctl0 = torch.tensor(s.ctl.astype(np.float64))
    L, g_ctl, g_s = _device_grad(m, pos, target, wt, ctl0, steps, every, 4)
    ref = FullRef(s, torch.float64)
    plan = ref.sample_plan(pos_h)
    tgt = target.cpu().double()
    wt64 = wt.cpu().double()
    shp = (NSTATE, s.nz, s.ny, s.nx)
    # float64 vector-Jacobian products chained along the device's states
    c32 = ctl0.numpy().astype(np.float32)
    sts = [m.new_state()]
    nst = torch.empty_like(sts[0])
    samples = []
    gpd = torch.empty((pos.shape[0], 4), dtype=torch.int32, device="cuda")
    for t in range(steps):
        m.step(sts[-1], nst, c32)
        sts.append(nst.clone())
        if (t + 1) % every == 0:
            o = torch.empty((pos.shape[0], len(FULL_SAMPLED)), device="cuda")
            m.sample(sts[-1], pos, o, gpd)
            samples.append(o)
    res = 2.0 * (torch.stack(samples).cpu().double() - tgt) * wt64 * wt64

    def chain(dtype):
        """Vector-Jacobian products chained back along the device's states:
        float64 linearised at the float32 values (Float32Values), or
        torch's float32 autograd."""
        r = FullRef(s, dtype)
        lam = [torch.zeros(shp[1:], dtype=dtype) for _ in range(NSTATE)]
        gc = torch.zeros(len(FULL_CTL_NAMES), dtype=torch.float64)
        cc = torch.tensor(c32.astype(np.float64), dtype=dtype, requires_grad=True)
        for t in reversed(range(steps)):
            if (t + 1) % every == 0:
                xs = [a.requires_grad_(True)
                      for a in r.state(sts[t + 1].cpu().numpy().reshape(shp))]
                sm = r.sample_full(xs, plan, FULL_SAMPLED)
                gs = torch.autograd.grad((sm * res[(t + 1) // every - 1].to(dtype)).sum(),
                                         [xs[i] for i in FULL_SAMPLED])
                for i, gg in zip(FULL_SAMPLED, gs):
                    lam[i] = lam[i] + gg
            st_h = sts[t].cpu().numpy().reshape(shp)
            xs = [a.requires_grad_(True) for a in r.state(st_h)]
            codes = _branches(s, st_h, c32)
            if dtype == torch.float64:
                with Float32Values():
                    y = r.step(xs, _ctl_list(cc, dtype), codes=codes)
            else:
                y = r.step(xs, _ctl_list(cc, dtype), codes=codes)
            g1, *gl = torch.autograd.grad(sum((a * b).sum() for a, b in zip(y, lam)),
                                          [cc, *xs], allow_unused=True)
            gc += g1.double()
            lam = [torch.zeros(shp[1:], dtype=dtype) if a is None else a for a in gl]
        return gc.numpy(), np.stack([a.double().numpy().reshape(-1) for a in lam])

    gc64, lam64 = chain(torch.float64)
    gc32, lam32 = chain(torch.float32)
    gsr = g_s.reshape(NSTATE, -1)

    def errs(gs, gc):
        es = {k: _rel(gs[list(v)], lam64[list(v)]) for k, v in GROUPS.items()
              if np.linalg.norm(lam64[list(v)]) > 0}
        ec = {FULL_CTL_NAMES[i]: abs(gc[i] - gc64[i]) / abs(gc64[i])
              for i in range(len(FULL_CTL_NAMES)) if gc64[i] != 0.0}
        return es, ec

    es, ec = errs(gsr, g_ctl)
    es32, ec32 = errs(lam32, gc32)
    _say("device vs chained float64 vjps: state", {k: f"{v:.1e}" for k, v in es.items()})
    _say("  ctl", {k: f"{v:.1e}" for k, v in ec.items()})
    _say("float32 autograd chain: state", {k: f"{v:.1e}" for k, v in es32.items()})
    _say("  ctl", {k: f"{v:.1e}" for k, v in ec32.items()})
    # bar: 1e-5, or 3x the float32 autograd chain's error (the b cancellation
    # of the step test compounds over the steps)
    bad = {k: v for k, v in es.items() if v > max(1e-5, 3.0 * es32[k])}
    bad.update({k: v for k, v in ec.items() if v > max(1e-5, 3.0 * ec32[k])})
    assert not bad, bad

    # float64 central differences of the restatement in each control, and
    # its float64 autograd gradient.  The step (1e-6 of the control's scale)
    # stays inside one branch of every limiter.
    def loss(c):
        out = ref.rollout_full(_ctl_list(c, torch.float64), ref.state(s.state), steps, every,
                               plan, FULL_SAMPLED)
        return (((out - tgt) * wt64) ** 2).sum()

    c64 = ctl0.clone().requires_grad_(True)
    (g64,) = torch.autograd.grad(loss(c64), [c64])
    g64 = g64.numpy()
    with torch.no_grad():
        for i in range(len(FULL_CTL_NAMES)):
            h = 1e-6 * max(abs(float(ctl0[i])), 0.1)
            cp, cm = ctl0.clone(), ctl0.clone()
            cp[i] += h
            cm[i] -= h
            fd = (float(loss(cp)) - float(loss(cm))) / (2 * h)
            # the float32 model (the reference's, which the device
            # differentiates exactly) and the float64 model linearise at
            # values that differ by float32 rounding; where a control's
            # gradient is a cancelling sum that moves it by up to ~3e-3
            model = abs(gc64[i] - g64[i])
            _say(f"control {FULL_CTL_NAMES[i]}: device {g_ctl[i]:.9g} fd {fd:.9g} "
                 f"float64 autograd {g64[i]:.9g} float32-valued float64 {gc64[i]:.9g} "
                 f"| device-fd {abs(g_ctl[i] - fd) / abs(fd):.1e} "
                 f"models {model / abs(fd):.1e}")
            assert abs(g64[i] - fd) <= 2e-4 * abs(fd), (FULL_CTL_NAMES[i], g64[i], fd)
            assert model <= 5e-3 * abs(fd), (FULL_CTL_NAMES[i], gc64[i], g64[i])
            assert abs(g_ctl[i] - fd) <= 1e-3 * abs(fd) + 2.0 * model, \
                (FULL_CTL_NAMES[i], g_ctl[i], fd)
    m.close()


def test_full_gradient_independent_of_checkpoint_interval():
    _say("full model checkpoint intervals")
    steps, every = 8, 4
    s, m, pos_h, pos, target, wt = _misfit_case(steps=steps, every=every)
    ctl0 = torch.tensor(s.ctl.astype(np.float64))
    res = [_device_grad(m, pos, target, wt, ctl0, steps, every, k) for k in (1, 3, 8)]
    L0, c0, g0 = res[0]
    for L, c, g in res[1:]:
        assert L == L0
        assert np.allclose(c, c0, rtol=1e-5, atol=0)
        assert _rel(g, g0) < 1e-5
    m.close()


def test_full_adjoint_cost_ratio():
    _say("full adjoint cost")
    steps, every = 20, 5
    s, m, pos_h, pos, target, wt = _misfit_case(dims=MID, steps=steps, every=every)
    ctl0 = torch.tensor(s.ctl.astype(np.float64))
    _device_grad(m, pos, target, wt, ctl0, steps, every, 10)       # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with torch.no_grad():
        full_rollout(ctl0, m.new_state(), m, steps, every, pos)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    _device_grad(m, pos, target, wt, ctl0, steps, every, 10)
    t2 = time.perf_counter()
    fwd, both = t1 - t0, t2 - t1
    _say(f"forward {fwd * 1e3:.2f} ms, forward+backward {both * 1e3:.2f} ms, "
         f"backward/forward {(both - fwd) / fwd:.2f}")
    assert np.isfinite(both)
    m.close()
