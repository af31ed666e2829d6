"""Host logic of the differentiation configuration (g3m3_amd/misfit.py): craft
placement, the loss, the mapping of theta into the workload and the
central-difference bookkeeping.  No GPU: the forward model is replaced by a
closed-form stand-in where one is needed."""
import numpy as np

from g3m3_amd import misfit as M


def test_craft_positions_in_shell_and_seeded():
    cfg = M.MisfitConfig()
    wl = M.workload_for(M.THETA0, cfg)
    p = M.craft_positions(cfg, wl)
    assert p.shape == (cfg.ncraft, 3) and p.dtype == np.float32
    r = np.linalg.norm(p.astype(np.float64), axis=1)   # planet radii
    assert np.all(r >= cfg.r_min * (1 - 1e-6)) and np.all(r <= cfg.r_max * (1 + 1e-6))
    assert np.array_equal(p, M.craft_positions(cfg, wl))
    assert not np.array_equal(p, M.craft_positions(M.MisfitConfig(seed=cfg.seed + 1), wl))


def test_craft_land_inside_a_box():
    """Each craft, placed by the reference's findgrid (which divides by
    re_equiv), falls inside a box, away from its faces, with trilinear
    weights in [0, 1]."""
    from oracle import restate as O
    cfg = M.MisfitConfig()
    wl = M.workload_for(M.THETA0, cfg)
    g = wl.grid
    mn = np.stack([g["grd_xmin"], g["grd_ymin"], g["grd_zmin"]], 1).astype(np.float32)
    mx = np.stack([g["grd_xmax"], g["grd_ymax"], g["grd_zmax"]], 1).astype(np.float32)
    xs = np.asarray(g["xspac"], np.float32)
    re = np.float32(wl.scal["re_equiv"])
    n = (wl.nx, wl.ny, wl.nz)
    for p in M.craft_positions(cfg, wl):
        gp = np.zeros(4, np.int32)
        O.findgrid(p, cfg.ngrd, mn, mx, xs, gp, re)
        b = int(gp[3]) - 1
        xa = p / re
        assert np.all(xa > mn[b]) and np.all(xa < mx[b]), p      # found, not clamped
        assert all(1 <= gp[a] <= n[a] - 1 for a in range(3)), gp
        cv = O.aux_cube(gp, mn, xs, re)
        w = (p - cv[0]) / (cv[1] - cv[0])
        assert np.all(w >= 0.0) and np.all(w <= 1.0), (p, w)


def test_loss():
    a = np.arange(24, dtype=np.float32).reshape(2, 3, 4)
    assert M.loss(a, a) == 0.0
    b = a.copy()
    b[1, 2, 3] += 2.0
    b[0, 0, 0] -= 1.0
    assert M.loss(b, a) == 5.0


def test_workload_for_maps_theta():
    cfg = M.MisfitConfig(L=8, ngrd=2)
    s0 = M.workload_for(M.THETA0, cfg).scal
    s = M.workload_for([1.0, -2.0, -6.0, 7.0, 450.0, 2e-3], cfg).scal
    assert s0["sbx_wind"] == 0.0 and s0["sby_wind"] == 0.0
    assert np.isclose(s["sbx_wind"] / s0["sbz_wind"], 1.0 / -5.0, rtol=1e-6)
    assert np.isclose(s["sby_wind"] / s0["sbz_wind"], -2.0 / -5.0, rtol=1e-6)
    assert np.isclose(s["sbz_wind"] / s0["sbz_wind"], 6.0 / 5.0, rtol=1e-6)
    assert np.isclose(s["srho"] / s0["srho"], 7.0 / 5.0, rtol=1e-6)
    assert np.isclose((s["spx"] / s["srho"]) / (s0["spx"] / s0["srho"]), 450.0 / 400.0,
                      rtol=1e-5)
    assert np.isclose(float(s["eta"]), 2e-3, rtol=1e-6)


def _quadratic_forward(monkeypatch, calls):
    """forward_many stand-in: a run's samples are theta itself, so the loss
    against a zero target is sum(theta^2) and its gradient 2 theta."""
    def fake(thetas, cfg=None, streams=8):
        calls.append((len(thetas), streams))
        out = [np.asarray(t, dtype=np.float64).copy() for t in thetas]
        return (out[0] if len(out) == 1 else out), 0.25 * len(thetas)
    monkeypatch.setattr(M, "forward_many", fake)


def test_gradient_fd_central_differences(monkeypatch):
    # moderate values: the loss difference of two sums cancels, with an
    # error of about |loss| * 1e-16 / h
    theta = np.array([0.0, 3.0, -5.0, 5.0, 40.0, 0.5])
    target = np.zeros(6)
    for together in (True, False):
        calls = []
        _quadratic_forward(monkeypatch, calls)
        g, secs = M.gradient_fd(theta, target, M.MisfitConfig(), together=together)
        # central differences are exact on a quadratic up to rounding;
        # theta_i = 0 is stepped by rel itself
        assert np.allclose(g, 2.0 * theta, rtol=1e-7, atol=1e-9)
        assert secs == 0.25 * 12
        if together:
            assert calls == [(12, 8)]              # every run stepped together
        else:
            assert calls == [(1, 1)] * 12          # one run at a time
