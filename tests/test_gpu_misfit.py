"""The differentiation configuration's forward model, loss and
central-difference gradient (g3m3_amd.misfit), on a shortened run: the
samples repeat bit for bit, the loss of a run against itself is 0, the craft
see non-zero fields, and the gradient is finite and repeats."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_misfit_forward_loss_gradient():
    from g3m3_amd import misfit as M
    cfg = M.MisfitConfig(steps=10, every=5)
    theta = M.THETA0.copy()
    print("misfit forward at theta0", flush=True)
    s1, _ = M.forward(theta, cfg)
    s2, _ = M.forward(theta, cfg)
    assert s1.shape == (2, cfg.ncraft, 4)
    assert np.array_equal(s1, s2)
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
    assert np.array_equal(g, g2)     # runs stepped together equal runs alone
