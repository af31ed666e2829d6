"""Pins the NumPy oracle to the reference's own Fortran, compiled unchanged
(serial, -ffp-contract=off) into oracle/_ref: every hot-path subroutine is
run on the same seeded inputs by both, and every array they can touch must
agree BIT FOR BIT.  Skipped where the reference build is absent (the GPU
box); the golden fixtures in tests/golden carry the same pinning there."""
import numpy as np
import pytest

from oracle import ref as R
from oracle import restate as O
from common import (FIELDS, World, push_ion_names, set_rho_names, set_speed_names,
                    bndry_outer_names, bndry_inner_names, lap_plasma_names,
                    flux_correct_names, flanks_names, corer_names, GRID)

pytestmark = pytest.mark.skipif(not R.available(), reason="oracle/_ref not built")

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = R.RefLib(omp=False)
    return _LIB


def run_both(w, name, names, pyname=None, **kw):
    wr, wo = w.copy(), w.copy()
    lib().call(name, *wr.args(names))
    getattr(O, pyname or name)(*wo.args(names), **kw)
    return wr, wo


def assert_changed(w0, w1, keys):
    """Guards against a vacuous pass: the call must have written these."""
    for k in keys:
        assert not np.array_equal(w0.a[k], w1.a[k]), f"{k} unchanged"


def assert_same(wr, wo, keys=None):
    bad = []
    for k in (keys or wr.a.keys()):
        a, b = wr.a[k], wo.a[k]
        if not np.array_equal(a, b, equal_nan=True):
            d = np.abs(a.astype(np.float64) - b.astype(np.float64))
            bad.append((k, float(d.max()), int((d > 0).sum())))
    assert not bad, f"oracle differs from reference: {bad[:8]}"


@pytest.fixture(params=[(13, 11, 9), (9, 12, 7)])
def world(request):
    nx, ny, nz = request.param
    w = World(nx, ny, nz, 2, seed=sum(request.param))
    w.s.update(box=2, delt=np.float32(0.037), vlim=np.float32(1.2), rx=np.float32(2.),
               ry=np.float32(2.), rz=np.float32(2.))
    return w


def test_calcur(world):
    names = ["bx", "by", "bz", "nx", "ny", "nz", "n_grids", "box", "curx", "cury", "curz",
             "rx", "ry", "rz"]
    assert_same(*run_both(world, "calcur", names))


def test_fnd_evel(world):
    names = ["qpx", "qpy", "qpz", "qrho", "hpx", "hpy", "hpz", "hrho", "opx", "opy", "opz",
             "orho", "curx", "cury", "curz", "evx", "evy", "evz", "tvx", "tvy", "tvz",
             "nx", "ny", "nz", "n_grids", "box", "rmassq", "rmassh", "rmasso", "reynolds"]
    assert_same(*run_both(world, "fnd_evel", names))


@pytest.mark.parametrize("box", [1, 2])
def test_BandE(world, box):
    world.s["box"] = box
    names = ["efldx", "efldy", "efldz", "bsx", "bsy", "bsz", "curx", "cury", "curz",
             "evx", "evy", "evz", "btot", "epres", "qrho", "hrho", "orho", "rst", "resist",
             "reynolds", "nx", "ny", "nz", "n_grids", "box", "rmassq", "rmassh", "rmasso",
             "ijmid", "nummid", "ijzero", "mbndry", "numzero", "mmid", "mzero",
             "rx", "ry", "rz"]
    assert_same(*run_both(world, "BandE", names))


def test_push_elec(world):
    names = ["wrkepres", "oldepres", "epres", "evx", "evy", "evz", "gamma", "gamma1",
             "nx", "ny", "nz", "n_grids", "box", "delt", "rx", "ry", "rz"]
    assert_same(*run_both(world, "push_elec", names))


def test_push_bfld(world):
    names = ["bx", "by", "bz", "oldbx", "oldby", "oldbz", "efldx", "efldy", "efldz",
             "nx", "ny", "nz", "n_grids", "box", "delt", "rx", "ry", "rz"]
    assert_same(*run_both(world, "push_bfld", names))


@pytest.mark.parametrize("sp,rm", [("q", "rmassq"), ("o", "rmasso")])
def test_push_ion(world, sp, rm):
    # strong field in part of the box so both skin-factor branches run
    world.a["bsx"][:, :, : world.nx // 2] *= 40
    world.a["btot"][...] = np.sqrt(world.a["bsx"] ** 2 + world.a["bsy"] ** 2
                                   + world.a["bsz"] ** 2)
    assert_same(*run_both(world, "push_ion", push_ion_names("", "old", "wrk", sp, rm)))


def test_set_rho(world):
    r = world.rng
    for f in ("qrho", "hrho", "orho"):
        m = r.random(world.a[f].shape) < 0.1
        world.a[f][m] = r.uniform(-1e-4, 1e-4, m.sum())
    for f in ("qpresx", "hpresy", "opresz", "epres"):
        m = r.random(world.a[f].shape) < 0.1
        world.a[f][m] *= -1
    wr, wo = run_both(world, "set_rho", set_rho_names(""))
    assert_same(wr, wo)
    assert_changed(world, wr, ["qrho", "hpresx", "opresxy", "epres"])


def test_set_speed_agrd(world):
    r = world.rng
    for f in ("qpx", "hpy", "opz"):
        m = r.random(world.a[f].shape) < 0.15
        world.a[f][m] *= 20
    m = r.random(world.a["epres"].shape) < 0.2
    world.a["epres"][m] *= 30
    world.s["vlim"] = np.float32(0.6 * 2)
    wr, wo = run_both(world, "set_speed_agrd", set_speed_names(""))
    assert_same(wr, wo)
    assert_changed(world, wr, ["qrho", "hpresxy", "opz", "epres"])
    assert wr.a["fastest"][0] > 0


@pytest.mark.parametrize("sp", ["q", "h"])
def test_lap_plasma(world, sp):
    assert_same(*run_both(world, "lap_plasma", lap_plasma_names(sp, "", "wrk")))


def test_lap_elec_bfld(world):
    names = ["epres", "wrkepres", "vvx", "vvy", "vvz", "nx", "ny", "nz", "n_grids", "box",
             "chirho", "chipxyz", "chierg", "delt"]
    assert_same(*run_both(world, "lap_elec", names))
    names = ["bx", "by", "bz", "wrkbx", "wrkby", "wrkbz", "vvx", "vvy", "vvz", "nx", "ny",
             "nz", "n_grids", "box", "chirho", "chipxyz", "chierg", "delt"]
    world.s["chierg"] = np.float32(0.7)
    assert_same(*run_both(world, "lap_bfld", names))


def test_fnd_vtot(world):
    names = ["qpx", "qpy", "qpz", "qrho", "hpx", "hpy", "hpz", "hrho", "opx", "opy", "opz",
             "orho", "vvx", "vvy", "vvz", "nx", "ny", "nz", "n_grids", "box",
             "rmassq", "rmassh", "rmasso"]
    assert_same(*run_both(world, "fnd_vtot", names))


def test_fcsmooth(world):
    names = ["qrho", "oldqrho", "wrkqrho", "nx", "ny", "nz", "n_grids", "box",
             np.float32(0.3), "vvx", "vvy", "vvz"]
    assert_same(*run_both(world, "fcsmooth", names))


def test_flux_correct(world):
    assert_same(*run_both(world, "flux_correct", flux_correct_names()))


def test_bndry_outer(world):
    world.s.update(srho=np.float32(25.), spress=np.float32(0.02395), spx=np.float32(10.),
                   spy=np.float32(0.), spz=np.float32(-0.1), sbx_wind=np.float32(0.),
                   sby_wind=np.float32(0.01), sbz_wind=np.float32(-0.244))
    assert_same(*run_both(world, "bndry_outer", bndry_outer_names("")))


def test_bndry_inner(world):
    world.s.update(erho=np.float32(1.), epress=np.float32(1.), alpha_e=np.float32(6.),
                   r_inner=np.float32(3.), sbx_wind=np.float32(0.), spress=np.float32(0.02))
    rot = lib().rotation()
    rot.v_rot, rot.r_rot = 0.3, 2.0
    assert_same(*run_both(world, "bndry_inner", bndry_inner_names(""), v_rot=0.3, r_rot=2.0))


@pytest.fixture
def nested():
    w = World(13, 13, 7, 3, seed=7, lists=False)
    w.a["t_old"] = np.array([0.0, 0.1, 0.1], dtype=np.float64)
    w.a["t_new"] = np.array([0.05, 0.3, 0.5], dtype=np.float64)
    w.s["t"] = R.F64(0.17)
    return w


@pytest.mark.parametrize("box", [1, 2])
def test_bndry_flanks(nested, box):
    nested.s["box"] = box
    assert_same(*run_both(nested, "bndry_flanks", flanks_names("wrk", "", "old")))


def test_bndry_flanks_aliased(nested):
    """The reference's 4th call passes the state as target and source
    (code/3d_main.f90:2945-2961)."""
    nested.s["box"] = 1
    assert_same(*run_both(nested, "bndry_flanks", flanks_names("", "", "old")))


@pytest.mark.parametrize("box", [1, 2])
def test_bndry_corer(nested, box):
    nested.s["box"] = box
    assert_same(*run_both(nested, "bndry_corer", corer_names("")))


@pytest.mark.parametrize("box", [1, 2])
def test_flanks_synced(nested, box):
    nested.s["box"] = box
    assert_same(*run_both(nested, "flanks_synced",
                          ["bx", "nx", "ny", "nz", "n_grids", "box"] + GRID))


def test_banded_tpose():
    nx, ny, nz = 7, 6, 5
    npts = nx * ny * nz
    M = O.divb_matrix(nx, ny, nz)
    pos = np.array([nx * ny, nx, 1, 0, -1, -nx, -nx * ny], np.int32)
    x = np.random.default_rng(3).uniform(-1, 1, npts).astype(np.float32)
    y1 = np.zeros(npts, np.float32)
    y2 = np.zeros(npts, np.float32)
    lib().call("banded_tpose", 7, npts, pos, M, x, y1)
    O.banded_tpose(7, npts, pos, M, x, y2)
    assert np.array_equal(y1, y2)


@pytest.mark.parametrize("errlim", [1e-4, 1e-6])
def test_sparse_solver(errlim):
    nx, ny, nz = 9, 8, 7
    npts = nx * ny * nz
    M = O.divb_matrix(nx, ny, nz)
    pos = np.array([nx * ny, nx, 1, 0, -1, -nx, -nx * ny], np.int32)
    rhs = np.random.default_rng(5).uniform(-1, 1, npts).astype(np.float32)
    s1 = np.zeros(npts, np.float32)
    s2 = np.zeros(npts, np.float32)
    r1 = np.zeros(1, np.float32)
    r2 = np.zeros(1, np.float32)
    lib().call("sparse_solver", 7, npts, pos, np.float32(errlim), False, 0, rhs, s1, r1, M)
    O.sparse_solver(7, npts, pos, np.float32(errlim), False, 0, rhs, s2, r2, M)
    assert np.array_equal(s1, s2) and r1[0] == r2[0]


def test_trilin_interp():
    rng = np.random.default_rng(9)
    for _ in range(20):
        g = np.sort(rng.uniform(-5, 5, (2, 3)), axis=0).astype(np.float32)
        s = (g[0] + rng.uniform(0, 1, 3) * (g[1] - g[0])).astype(np.float32)
        q = rng.uniform(-1, 1, (2, 2, 2)).astype(np.float32)
        a, b = np.zeros(1, np.float32), np.zeros(1, np.float32)
        lib().call("trilin_interp", s, g, q, a)
        O.trilin_interp(s, g, q, b)
        assert a[0] == b[0]


def test_fnd_pres(world):
    w = world.copy()
    shp = w.a["qrho"].shape
    for k in ("p_para", "p_perp", "p_cross"):
        w.a[k] = np.zeros(shp, np.float32)
    names = ["qpresx", "qpresy", "qpresz", "qpresxy", "qpresxz", "qpresyz", "p_para", "p_perp",
             "p_cross", "vvx", "vvy", "vvz", "bsx", "bsy", "bsz", "nx", "ny", "nz", "n_grids",
             "box"]
    wr, wo = run_both(w, "fnd_pres", names)
    assert_changed(w, wr, ["p_para", "p_perp", "p_cross"])
    assert_same(wr, wo)


def _grid_extents(ngrd, L=6):
    xs = np.array([2.0 ** b for b in range(ngrd)], np.float32)
    mn = np.stack([-L * xs, -L * xs, -L / 2 * xs], axis=1).astype(np.float32)
    mx = np.stack([L * xs, L * xs, L / 2 * xs], axis=1).astype(np.float32)
    return xs, np.ascontiguousarray(mn), np.ascontiguousarray(mx)


def test_findgrid():
    """Inside the smallest box, inside only outer boxes, on a box face
    (strict inequalities), and outside every box (clamped to the outermost,
    with the reference's message)."""
    ngrd, re = 3, np.float32(0.5)
    xs, mn, mx = _grid_extents(ngrd)
    rng = np.random.default_rng(5)
    pts = [rng.uniform(-30, 30, 3) for _ in range(40)]
    pts += [np.array([6 * 0.5, 0.0, 0.0]), np.array([100.0, -100.0, 3.0]),
            np.array([0.0, 0.0, -200.0])]
    for p in pts:
        p = np.asarray(p, np.float32)
        a, b = np.zeros(4, np.int32), np.zeros(4, np.int32)
        lib().call("findgrid", p, ngrd, mn, mx, xs, a, re, R.Chars("craft01", 8))
        O.findgrid(p, ngrd, mn, mx, xs, b, re)
        assert np.array_equal(a, b), (p, a, b)


# --- initialisation: dipole, mak_dip_grd, set_resist --------------------------

ROT = dict(rot_angle=0.7, xdip=0.00001, ydip=-0.3, zdip=0.2,
           sin_tilt=float(np.float32(np.sin(0.35))), cos_tilt=float(np.float32(np.cos(0.35))),
           b0=270.0)


def _set_rotation(**kw):
    rot = lib().rotation()
    for k, v in kw.items():
        setattr(rot, k, v)


def test_dipole_points():
    """The dipole at 2000 points, rotated and tilted: bit for bit."""
    _set_rotation(**ROT)
    rng = np.random.default_rng(41)
    pts = rng.uniform(-12, 12, (2000, 3)).astype(np.float32)
    for p in pts:
        r = [np.zeros(1, np.float32) for _ in range(3)]
        o = [np.zeros(1, np.float32) for _ in range(3)]
        lib().call("dipole", *r, p[0], p[1], p[2])
        O.dipole(*o, p[0], p[1], p[2], **ROT)
        assert [x[0] for x in r] == [x[0] for x in o], p


@pytest.mark.parametrize("box", [1, 2])
def test_mak_dip_grd(box):
    _set_rotation(**ROT)
    nx, ny, nz, ng, mb = 13, 11, 9, 2, 1
    w = World(nx, ny, nz, ng, seed=3, lists=False)
    shp = (ng, nz, ny, nx)
    ij = np.zeros((4, 3, mb), np.int32)
    ij[:, :, 0] = [[7, 6, 5], [1, 1, 1], [13, 11, 9], [2, 3, 4]]
    num = np.array([3], np.int32)
    outs = []
    for call in (lambda *a: lib().call("mak_dip_grd", *a), lambda *a: O.mak_dip_grd(*a, **ROT)):
        b = [np.full(shp, 5.0, np.float32) for _ in range(3)]
        call(*b, nx, ny, nz, ng, mb, box, ij, num, 4, np.float32(3.0),
             w.a["grd_xmin"], w.a["grd_xmax"], w.a["grd_ymin"], w.a["grd_ymax"],
             w.a["grd_zmin"], w.a["grd_zmax"])
        outs.append(b)
    for r, o in zip(*outs):
        assert np.array_equal(r, o)
    assert np.abs(outs[0][2][box - 1]).max() > 0
    if box == 1:   # the first numzero listed points are zeroed, the fourth is not
        assert outs[0][0][0, 4, 5, 6] == 0.0 and outs[0][0][0, 3, 2, 1] != 0.0


def test_set_resist():
    nx, ny, nz, mb = 9, 8, 7, 2
    rng = np.random.default_rng(8)

    def lst(n, m):
        a = np.zeros((m, 3, mb), np.int32)
        a[:, 0, :] = rng.integers(1, nx + 1, (m, mb))
        a[:, 1, :] = rng.integers(1, ny + 1, (m, mb))
        a[:, 2, :] = rng.integers(1, nz + 1, (m, mb))
        return a

    ijz, ijm, ijs = lst(0, 6), lst(0, 9), lst(0, 12)
    nz_, nm_, ns_ = (np.array([2, 5], np.int32), np.array([9, 4], np.int32),
                     np.array([1, 11], np.int32))
    outs = []
    for fn in (lambda *a: lib().call("set_resist", *a), O.set_resist):
        rst = np.full((mb, nz, ny, nx), 3.0, np.float32)
        fn(rst, nx, ny, nz, mb, np.float32(64.0), ijz, nz_, ijm, nm_, ijs, ns_, 12, 9, 6,
           np.float32(1.0))
        outs.append(rst)
    assert np.array_equal(*outs)
    assert (outs[0] != 0).sum() > 10


def test_init_plasma_field_and_resist_vs_reference():
    """The oracle's initialisation block against the reference's own pieces:
    b0 of every box equals the reference dipole subroutine called as the
    driver calls it (at positions already shifted by the dipole centre,
    code/3d_main.f90:1256-1274) wherever r >= zero_bndry, and the
    resistivity equals the reference set_resist on the lists produced."""
    from common import init_case, init_oracle
    c = init_case()
    d, _ = init_oracle(c)
    s = d["s"]
    _set_rotation(**{k: s[k] for k in ("rot_angle", "xdip", "ydip", "zdip", "sin_tilt",
                                        "cos_tilt", "b0")})
    F = np.float32
    nx, ny, nz = d["nx"], d["ny"], d["nz"]
    rng = np.random.default_rng(2)
    sr, cr = O._sinf(s["rot_angle"]), O._cosf(s["rot_angle"])
    checked = 0
    for b in range(d["ngrd"]):
        g0, g1 = d["gmin"][b], d["gmax"][b]
        dx, dy, dz = ((g1[a] - g0[a]) / (F(n) - F(1.)) for a, n in ((0, nx), (1, ny), (2, nz)))
        for _ in range(300):
            i, j, k = rng.integers(0, nx), rng.integers(0, ny), rng.integers(0, nz)
            ax = g0[0] + dx * F(i) - F(s["xdip"])
            ay = g0[1] + dy * F(j) - F(s["ydip"])
            az = g0[2] + dz * F(k) - F(s["zdip"])
            xr = ax * cr + ay * sr
            yr = -(ax * sr) + ay * cr
            xp = xr * F(s["cos_tilt"]) - az * F(s["sin_tilt"])
            zp = xr * F(s["sin_tilt"]) + az * F(s["cos_tilt"])
            ar = np.sqrt(xp * xp + yr * yr + zp * zp)
            r = [np.zeros(1, F) for _ in range(3)]
            lib().call("dipole", *r, ax, ay, az)
            got = [d["b0"][q][b, k, j, i] for q in range(3)]
            if ar < F(s["zero_bndry"]):
                assert got == [0.0, 0.0, 0.0]
            else:
                assert got == [x[0] for x in r]
                checked += 1
    assert checked > 400
    L = d["lists"]
    rst = np.zeros_like(d["rst"])
    lib().call("set_resist", rst, nx, ny, nz, 1, F(s["resist"]), L["ijzero"], L["numzero"],
               L["ijmid"], L["nummid"], L["ijsrf"], L["numsrf"], d["mcap"], d["mcap"], d["mcap"],
               F(1.0))
    assert np.array_equal(rst, d["rst"])
    assert L["numzero"][0] > 0 and L["nummid"][0] > 0 and L["numsrf"][0] > 0


# --- craft placement: linterp, find_aux_cube ------------------------------------

def test_linterp():
    rng = np.random.default_rng(9)
    for _ in range(300):
        a = rng.uniform(-20, 20, 4).astype(np.float32)
        b = rng.uniform(-20, 20, 4).astype(np.float32)
        if rng.uniform() < 0.1:
            b[3] = a[3]
        t = np.float32(rng.uniform(-20, 20))
        r, o = np.zeros(3, np.float32), np.zeros(3, np.float32)
        lib().call("linterp", a, b, t, r)
        O.linterp(a, b, t, o)
        assert np.array_equal(r, o), (a, b, t)


def test_find_aux_cube():
    """Between trajectory points, at and past the second point's time, and
    with both points at one time; cube vertices and gridpt bit for bit."""
    ngrd, re = 3, np.float32(0.5)
    xs, mn, mx = _grid_extents(ngrd)
    rng = np.random.default_rng(13)
    for q in range(120):
        x = np.append(rng.uniform(-12, 12, 3), rng.uniform(0, 10)).astype(np.float32)
        z = np.append(rng.uniform(-12, 12, 3), x[3] + rng.uniform(0.1, 5)).astype(np.float32)
        if q % 10 == 0:
            z[3] = x[3]
        ut = np.float32(x[3] + rng.uniform(-1, 6))
        outs = []
        for call in ("ref", "oracle"):
            nt = np.zeros(2, np.int32)
            rec = np.ones(1, np.int32)
            gp = np.zeros(4, np.int32)
            s = np.zeros(3, np.float32)
            cv = np.zeros((2, 3), np.float32)
            if call == "ref":
                lib().call("find_aux_cube", R.Chars("craft01", 8), 1, 40, ngrd, mn, mx, xs, re,
                           x, z, ut, nt, rec, gp, s, cv)
            else:
                O.find_aux_cube("craft01", 1, 40, ngrd, mn, mx, xs, re, x, z, ut, nt, rec, gp, s,
                                cv)
            outs.append((gp, s, cv, rec))
        for u, v in zip(*outs):
            assert np.array_equal(u, v), (q, outs)


@pytest.mark.parametrize("shift", [0.0, 0.25, -0.25])
def test_flux_counter(world, shift):
    """Wall sums in the reference's order and the in/out bookkeeping, with
    the h flux positive, negative and mixed on the walls."""
    for sp in "qho":
        for c in "xyz":
            world.a[sp + "p" + c] += np.float32(shift)
    names = ["qpx", "qpy", "qpz", "hpx", "hpy", "hpz", "opx", "opy", "opz", "nx", "ny", "nz",
             "n_grids", "box", "rmassq", "rmassh", "rmasso", "grid_spacing", "qflux_in",
             "qflux_out", "hflux_in", "hflux_out", "oflux_in", "oflux_out"]
    world.s["grid_spacing"] = np.float32(2.0)
    for k in names[18:]:
        world.a[k] = np.zeros(1, np.float32)
    wr, wo = run_both(world, "flux_counter", names)
    assert_same(wr, wo, names[18:])
    assert any(wr.a[k][0] != 0 for k in names[18:])
