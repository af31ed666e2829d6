"""The differentiation configuration's forward model, loss and
central-difference gradient (g3m3_amd.misfit), on a shortened run: the
samples repeat bit for bit, the loss of a run against itself is 0, the craft
see non-zero fields, and the gradient is finite and repeats."""
import numpy as np
import pytest

from common import same_bits

pytestmark = pytest.mark.gpu


def test_misfit_forward_loss_gradient():
    from g3m3_amd import misfit as M
    cfg = M.MisfitConfig(steps=10, every=5)
    theta = M.THETA0.copy()
    print("misfit forward at theta0", flush=True)
    s1, _ = M.forward(theta, cfg)
    s2, _ = M.forward(theta, cfg)
    assert s1.shape == (2, cfg.ncraft, 4)
    assert same_bits(s1, s2)
    assert M.loss(s1, s2) == 0.0
    assert (np.abs(s1[..., 3]) > 0).sum() >= cfg.ncraft   # densities seen
    target_theta = theta * np.array([1.0, 1.0, 1.2, 1.1, 1.05, 1.0]) + \
        np.array([1.0, -1.0, 0.0, 0.0, 0.0, 0.0])
    target, _ = M.forward(target_theta, cfg)
    assert M.loss(s1, target) > 0.0
    print("misfit gradient", flush=True)
    g, secs = M.gradient_fd(theta, target, cfg)
    assert np.all(np.isfinite(g)) and np.any(g != 0.0) and secs > 0
    g2, _ = M.gradient_fd(theta, target, cfg, together=False)
    assert same_bits(g, g2)     # runs stepped together equal runs alone


def test_craft_samples_equal_oracle():
    """After two full steps, the device's craft samples (findgrid, aux cube
    and trilin_interp of b and the three densities) equal the oracle's on
    the same state bit for bit; every craft lies inside a box (gridpt box
    >= 1, found rather than clamped) and each sampled density lies within
    the range of that box's field."""
    from g3m3_amd import misfit as M
    from g3m3_amd.driver import Batch
    from oracle import restate as O
    cfg = M.MisfitConfig(steps=2, every=2)
    wl = M.workload_for(M.THETA0, cfg)
    print("craft samples after 2 steps of 3 x 81x81x41", flush=True)
    batch = Batch([wl], streams=1)
    batch.init_dt()
    for _ in range(2):
        batch.full_step()
    run = batch.runs[0]
    sm = M._Sampler(run, wl, cfg)
    sm.sample()
    got, gp = sm.out.cpu().numpy(), sm.gp.cpu().numpy()
    st = run.state_host()
    pos = M.craft_positions(cfg, wl)
    want, wgp = O.sample_fields(pos, cfg.ngrd, sm.mn, sm.mx, sm.xs,
                                np.float32(wl.scal["re_equiv"]), [st[k] for k in sm.NAMES])
    assert np.all(gp[:, 3] >= 1) and np.all(gp[:, 0] >= 1), gp
    assert same_bits(gp, wgp)
    assert same_bits(got, want)
    for c in range(cfg.ncraft):
        b = gp[c, 3] - 1
        for q, k in enumerate(sm.NAMES[3:], start=3):
            f = st[k][b]
            assert f.min() <= got[c, q] <= f.max(), (c, k)
