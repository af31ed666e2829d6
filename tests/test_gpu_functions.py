"""GPU parity, one entry point at a time: every C-ABI function is run on the
device and compared with the NumPy oracle on the same seeded inputs.

Bar: bit-exact (the kernels keep the reference's single-precision
expression order with no fused multiply-add).  Where a result is
legitimately not bit-exact the tolerance is written in the test."""
import numpy as np
import pytest

from common import same_bits

from common import (World, push_ion_names, set_rho_names, set_speed_names, bndry_outer_names,
                    bndry_inner_names, lap_plasma_names, flux_correct_names, flanks_names,
                    corer_names, GRID)
from oracle import restate as O

pytestmark = pytest.mark.gpu

# arrays the C-ABI reads on the host
HOST_ARRAYS = set(GRID) | {"xspac", "numsrf", "nummid", "numzero", "t_old", "t_new",
                           "pxmax", "pymax", "pzmax", "csmax", "alfmax", "fastest"}


@pytest.fixture(scope="module")
def dev():
    import torch
    from g3m3_amd import _lib
    assert torch.cuda.is_available()
    _lib.lib()
    return torch


def to_dev(torch, w):
    d = {}
    for k, v in w.a.items():
        if k in HOST_ARRAYS or not isinstance(v, np.ndarray):
            d[k] = v
        else:
            d[k] = torch.from_numpy(np.ascontiguousarray(v)).cuda()
    return d


def gpu_call(torch, w, name, names, extra=()):
    from g3m3_amd import _lib
    d = to_dev(torch, w)
    args = [d[n] if isinstance(n, str) and n in d else (w.s[n] if isinstance(n, str) else n)
            for n in names]
    r = _lib.call(name, *args, *extra)
    _lib.sync()
    out = w.copy()
    for k, v in d.items():
        if hasattr(v, "cpu"):
            out.a[k] = v.cpu().numpy()
    return out, r


def oracle_call(w, name, names, **kw):
    o = w.copy()
    getattr(O, name)(*o.args(names), **kw)
    return o


def assert_bits(wg, wo, keys, tol=0.0):
    bad = []
    for k in keys:
        a, b = wg.a[k], wo.a[k]
        if tol == 0.0:
            ok = same_bits(a, b)
        else:
            sc = max(float(np.abs(b).max()), 1e-30)
            ok = float(np.abs(a.astype(np.float64) - b).max()) <= tol * sc
        if not ok:
            d = np.abs(a.astype(np.float64) - b.astype(np.float64))
            bad.append((k, float(d.max()), int((d > 0).sum())))
    assert not bad, f"GPU differs from oracle: {bad[:8]}"


@pytest.fixture(params=[(13, 11, 9), (33, 17, 12)])
def world(request):
    nx, ny, nz = request.param
    w = World(nx, ny, nz, 2, seed=100 + sum(request.param))
    w.s.update(box=2, delt=np.float32(0.037), vlim=np.float32(1.2), rx=np.float32(2.),
               ry=np.float32(2.), rz=np.float32(2.), eta=np.float32(0.0))
    return w


def test_calcur(dev, world):
    names = ["bx", "by", "bz", "nx", "ny", "nz", "n_grids", "box", "curx", "cury", "curz",
             "rx", "ry", "rz"]
    wg, _ = gpu_call(dev, world, "calcur", names)
    assert_bits(wg, oracle_call(world, "calcur", names), ["curx", "cury", "curz"])


def test_fnd_evel(dev, world):
    names = ["qpx", "qpy", "qpz", "qrho", "hpx", "hpy", "hpz", "hrho", "opx", "opy", "opz",
             "orho", "curx", "cury", "curz", "evx", "evy", "evz", "tvx", "tvy", "tvz",
             "nx", "ny", "nz", "n_grids", "box", "rmassq", "rmassh", "rmasso", "reynolds"]
    wg, _ = gpu_call(dev, world, "fnd_evel", names)
    assert_bits(wg, oracle_call(world, "fnd_evel", names),
                ["evx", "evy", "evz", "tvx", "tvy", "tvz"])


@pytest.mark.parametrize("box", [1, 2])
def test_bande(dev, world, box):
    world.s["box"] = box
    names = ["efldx", "efldy", "efldz", "bsx", "bsy", "bsz", "curx", "cury", "curz",
             "evx", "evy", "evz", "btot", "epres", "qrho", "hrho", "orho", "rst", "resist",
             "reynolds", "nx", "ny", "nz", "n_grids", "box", "rmassq", "rmassh", "rmasso",
             "ijmid", "nummid", "ijzero", "mbndry", "numzero", "mmid", "mzero",
             "rx", "ry", "rz"]
    wg, _ = gpu_call(dev, world, "bande", names)
    assert_bits(wg, oracle_call(world, "BandE", names), ["efldx", "efldy", "efldz"])


def test_push_elec(dev, world):
    names = ["wrkepres", "oldepres", "epres", "evx", "evy", "evz", "gamma", "gamma1",
             "nx", "ny", "nz", "n_grids", "box", "delt", "rx", "ry", "rz"]
    wg, _ = gpu_call(dev, world, "push_elec", names)
    assert_bits(wg, oracle_call(world, "push_elec", names), ["epres"])


def test_push_bfld(dev, world):
    names = ["bx", "by", "bz", "oldbx", "oldby", "oldbz", "efldx", "efldy", "efldz",
             "nx", "ny", "nz", "n_grids", "box", "delt", "rx", "ry", "rz"]
    wg, _ = gpu_call(dev, world, "push_bfld", names)
    assert_bits(wg, oracle_call(world, "push_bfld", names), ["bx", "by", "bz"])


@pytest.mark.parametrize("sp,rm", [("q", "rmassq"), ("o", "rmasso")])
def test_push_ion(dev, world, sp, rm):
    world.a["bsx"][:, :, : world.nx // 2] *= 40
    world.a["btot"][...] = np.sqrt(world.a["bsx"] ** 2 + world.a["bsy"] ** 2
                                   + world.a["bsz"] ** 2)
    names = push_ion_names("", "old", "wrk", sp, rm)
    wg, _ = gpu_call(dev, world, "push_ion", names)
    keys = [sp + f for f in ("rho", "presx", "presy", "presz", "px", "py", "pz", "presxy",
                             "presxz", "presyz")] + ["vvx", "vvy", "vvz"]
    assert_bits(wg, oracle_call(world, "push_ion", names), keys)


def test_set_rho(dev, world):
    r = world.rng
    for f in ("qrho", "hrho", "orho"):
        m = r.random(world.a[f].shape) < 0.1
        world.a[f][m] = r.uniform(-1e-4, 1e-4, m.sum())
    for f in ("qpresx", "hpresy", "opresz", "epres"):
        m = r.random(world.a[f].shape) < 0.1
        world.a[f][m] *= -1
    names = set_rho_names("")
    wg, _ = gpu_call(dev, world, "set_rho", names)
    keys = [k for k in names if isinstance(k, str) and k in world.a]
    assert_bits(wg, oracle_call(world, "set_rho", names), keys)


def test_set_speed_agrd(dev, world):
    r = world.rng
    for f in ("qpx", "hpy", "opz"):
        m = r.random(world.a[f].shape) < 0.15
        world.a[f][m] *= 20
    m = r.random(world.a["epres"].shape) < 0.2
    world.a["epres"][m] *= 30
    names = set_speed_names("")
    wg, _ = gpu_call(dev, world, "set_speed_agrd", names)
    wo = oracle_call(world, "set_speed_agrd", names)
    keys = [k for k in names if isinstance(k, str) and k in world.a]
    assert_bits(wg, wo, keys)
    assert wg.a["fastest"][0] > 0


@pytest.mark.parametrize("sp", ["q", "o"])
def test_lap_plasma(dev, world, sp):
    names = lap_plasma_names(sp, "", "wrk")
    wg, _ = gpu_call(dev, world, "lap_plasma", names)
    keys = ["wrk" + sp + f for f in ("rho", "px", "py", "pz", "presx", "presy", "presz",
                                     "presxy", "presxz", "presyz")]
    assert_bits(wg, oracle_call(world, "lap_plasma", names), keys)


def test_lap_elec_bfld_vtot(dev, world):
    names = ["epres", "wrkepres", "vvx", "vvy", "vvz", "nx", "ny", "nz", "n_grids", "box",
             "chirho", "chipxyz", "chierg", "delt"]
    wg, _ = gpu_call(dev, world, "lap_elec", names)
    assert_bits(wg, oracle_call(world, "lap_elec", names), ["wrkepres"])
    world.s["chierg"] = np.float32(0.7)
    names = ["bx", "by", "bz", "wrkbx", "wrkby", "wrkbz", "vvx", "vvy", "vvz", "nx", "ny",
             "nz", "n_grids", "box", "chirho", "chipxyz", "chierg", "delt"]
    wg, _ = gpu_call(dev, world, "lap_bfld", names)
    assert_bits(wg, oracle_call(world, "lap_bfld", names), ["wrkbx", "wrkby", "wrkbz"])
    names = ["qpx", "qpy", "qpz", "qrho", "hpx", "hpy", "hpz", "hrho", "opx", "opy", "opz",
             "orho", "vvx", "vvy", "vvz", "nx", "ny", "nz", "n_grids", "box",
             "rmassq", "rmassh", "rmasso"]
    wg, _ = gpu_call(dev, world, "fnd_vtot", names)
    assert_bits(wg, oracle_call(world, "fnd_vtot", names), ["vvx", "vvy", "vvz"])


def test_flux_correct(dev, world):
    names = flux_correct_names()
    wg, _ = gpu_call(dev, world, "flux_correct", names)
    wo = oracle_call(world, "flux_correct", names)
    keys = [k for k in names[:102] if not k.startswith(("wrk", "old"))] + \
           ["epres", "bx", "by", "bz"]
    assert_bits(wg, wo, keys)


@pytest.mark.parametrize("shape", [(70, 19, 37), (64, 8, 16), (127, 9, 18), (5, 5, 5)])
def test_flux_correct_tiles(dev, shape):
    """The one-pass FCT works on 63 x 8 tiles and runs of 16 planes: boxes
    that end one cell into a tile, fill tiles exactly, or span several in
    each direction, down to the smallest box."""
    w = World(*shape, 2, seed=7 + sum(shape))
    w.s.update(box=2)
    names = flux_correct_names()
    wg, _ = gpu_call(dev, w, "flux_correct", names)
    wo = oracle_call(w, "flux_correct", names)
    keys = [k for k in names[:102] if not k.startswith(("wrk", "old"))] + \
           ["epres", "bx", "by", "bz"]
    assert_bits(wg, wo, keys)


def test_bndry_outer(dev, world):
    world.s.update(srho=np.float32(25.), spress=np.float32(0.02395), spx=np.float32(10.),
                   spy=np.float32(0.5), spz=np.float32(-0.1), sbx_wind=np.float32(0.),
                   sby_wind=np.float32(0.01), sbz_wind=np.float32(-0.244))
    world.s["n_grids"] = 2
    names = bndry_outer_names("")
    wg, _ = gpu_call(dev, world, "bndry_outer", names)
    keys = [k for k in names if isinstance(k, str) and k in world.a and k[-1] != "0"]
    assert_bits(wg, oracle_call(world, "bndry_outer", names), keys)


def test_bndry_inner(dev, world):
    world.s.update(erho=np.float32(1.), epress=np.float32(1.), alpha_e=np.float32(6.),
                   r_inner=np.float32(3.), sbx_wind=np.float32(0.), spress=np.float32(0.02),
                   v_rot=np.float32(0.3), r_rot=np.float32(2.0))
    names = bndry_inner_names("") + ["v_rot", "r_rot"]
    wg, _ = gpu_call(dev, world, "bndry_inner", names)
    wo = world.copy()
    O.bndry_inner(*wo.args(names[:-2]), v_rot=0.3, r_rot=2.0)
    keys = [k for k in names[:40] if isinstance(k, str) and k in world.a and
            world.a[k].ndim == 4]
    assert_bits(wg, wo, keys)


@pytest.fixture
def nested():
    w = World(13, 13, 7, 3, seed=17, lists=False)
    w.a["t_old"] = np.array([0.0, 0.1, 0.1], dtype=np.float64)
    w.a["t_new"] = np.array([0.05, 0.3, 0.5], dtype=np.float64)
    return w


@pytest.mark.parametrize("box,tgt,src", [(1, "wrk", ""), (2, "wrk", ""), (1, "", "")])
def test_bndry_flanks(dev, nested, box, tgt, src):
    from g3m3_amd import _lib
    from g3m3_amd import signatures as sig
    nested.s["box"] = box
    t = 0.17
    d = to_dev(dev, nested)
    ptr = lambda pre: np.array([d[pre + k].data_ptr() for k in sig.FLANK_ORDER], np.uint64)
    _lib.call("bndry_flanks_v", ptr(tgt), ptr(src), ptr("old"), d["work"], nested.nx,
              nested.ny, nested.nz, nested.ngrd, box, nested.a["t_old"], nested.a["t_new"],
              _lib.F64(t), *[nested.a[g] for g in GRID])
    _lib.sync()
    wo = nested.copy()
    wo.s["t"] = t
    O.bndry_flanks(*wo.args(flanks_names(tgt, src, "old")))
    for k in sig.FLANK_ORDER:
        got = d[tgt + k].cpu().numpy()
        assert same_bits(got, wo.a[tgt + k]), k


@pytest.mark.parametrize("box", [1, 2])
def test_bndry_corer(dev, nested, box):
    from g3m3_amd import _lib
    from g3m3_amd import signatures as sig
    d = to_dev(dev, nested)
    ptrs = np.array([d[k].data_ptr() for k in sig.CORER_ORDER], np.uint64)
    _lib.call("bndry_corer_v", ptrs, nested.nx, nested.ny, nested.nz, nested.ngrd, box,
              *[nested.a[g] for g in GRID])
    _lib.sync()
    wo = nested.copy()
    wo.s["box"] = box
    O.bndry_corer(*wo.args(corer_names("")))
    for k in sig.CORER_ORDER:
        assert same_bits(d[k].cpu().numpy(), wo.a[k]), k


@pytest.mark.parametrize("box", [1, 2])
def test_flanks_synced(dev, nested, box):
    nested.s["box"] = box
    names = ["bx", "nx", "ny", "nz", "n_grids", "box"] + GRID
    wg, _ = gpu_call(dev, nested, "flanks_synced", names)
    assert_bits(wg, oracle_call(nested, "flanks_synced", names), ["bx"])


def test_trilin_interp(dev):
    from g3m3_amd import _lib
    rng = np.random.default_rng(9)
    for _ in range(10):
        g = np.sort(rng.uniform(-5, 5, (2, 3)), axis=0).astype(np.float32)
        s = (g[0] + rng.uniform(0, 1, 3) * (g[1] - g[0])).astype(np.float32)
        q = rng.uniform(-1, 1, (2, 2, 2)).astype(np.float32)
        a, b = np.zeros(1, np.float32), np.zeros(1, np.float32)
        _lib.call("trilin_interp", s, g, q, a)
        O.trilin_interp(s, g, q, b)
        assert a[0] == b[0]


def test_divb_solves_poisson(dev):
    """divB_correct: the matrix-free CGNR reaches the reference criterion and
    removes the divergence it solved for; compared with the oracle's solve
    of the same system (dots in double vs sequential single: agreement to
    the solver tolerance, not bitwise)."""
    from g3m3_amd import _lib
    import ctypes
    w = World(17, 15, 11, 1, seed=3, lists=False)
    errlim = np.float32(1e-4)
    d = to_dev(dev, w)
    it = ctypes.c_longlong(0)
    r2 = ctypes.c_double(0)
    st = _lib.call("divb_correct_ex", w.nx, w.ny, w.nz, 1, 1, w.a["xspac"], errlim,
                   d["bx"], d["by"], d["bz"], 1, ctypes.c_longlong(0), it, r2)
    _lib.check("divb")
    assert st == 1 and it.value > 0
    wo = w.copy()
    pos = np.array([w.nx * w.ny, w.nx, 1, 0, -1, -w.nx, -w.nx * w.ny], np.int32)
    O.divB_correct(w.nx, w.ny, w.nz, 1, 1, 7, pos, wo.a["xspac"], errlim,
                   wo.a["bx"], wo.a["by"], wo.a["bz"])
    for k in ("bx", "by", "bz"):
        got = d[k].cpu().numpy()
        ref = wo.a[k]
        assert np.abs(got - ref).max() <= 0.05 * np.abs(w.a[k]).max()


def test_fnd_pres(dev, world):
    w = world.copy()
    shp = w.a["qrho"].shape
    for k in ("p_para", "p_perp", "p_cross"):
        w.a[k] = np.zeros(shp, np.float32)
    names = ["qpresx", "qpresy", "qpresz", "qpresxy", "qpresxz", "qpresyz", "p_para", "p_perp",
             "p_cross", "vvx", "vvy", "vvz", "bsx", "bsy", "bsz", "nx", "ny", "nz", "n_grids",
             "box"]
    wg, _ = gpu_call(dev, w, "fnd_pres", names)
    wo = oracle_call(w, "fnd_pres", names)
    assert_bits(wg, wo, ["p_para", "p_perp", "p_cross"])


def test_findgrid_host(dev):
    """The host findgrid entry point against the oracle (inside, on a face,
    outside every box)."""
    from g3m3_amd import _lib
    import ctypes
    ngrd, re = 3, np.float32(0.5)
    xs = np.array([1.0, 2.0, 4.0], np.float32)
    mn = np.ascontiguousarray(np.stack([-6 * xs, -6 * xs, -3 * xs], 1).astype(np.float32))
    mx = np.ascontiguousarray(np.stack([6 * xs, 6 * xs, 3 * xs], 1).astype(np.float32))
    rng = np.random.default_rng(12)
    for p in list(rng.uniform(-30, 30, (30, 3))) + [np.array([3.0, 0, 0])]:
        p = np.asarray(p, np.float32)
        a, b = np.zeros(4, np.int32), np.zeros(4, np.int32)
        name = ctypes.create_string_buffer(b"craft01 ", 8)
        _lib.call("findgrid", p, ngrd, mn, mx, xs, a, re, name, _lib.ByValue(ctypes.c_size_t(8)))
        O.findgrid(p, ngrd, mn, mx, xs, b, re)
        assert same_bits(a, b), (p, a, b)


def test_sample_fields(dev):
    """Craft sampling of several fields at many points in one launch against
    the oracle's findgrid + aux cube + trilin_interp per point."""
    from g3m3_amd import _lib
    nx, ny, nz, ngrd = 13, 13, 7, 3
    re = np.float32(0.5)
    xs = np.array([1.0, 2.0, 4.0], np.float32)
    mn = np.ascontiguousarray(np.stack([-6 * xs, -6 * xs, -3 * xs], 1).astype(np.float32))
    mx = np.ascontiguousarray(np.stack([6 * xs, 6 * xs, 3 * xs], 1).astype(np.float32))
    rng = np.random.default_rng(21)
    fields = [rng.uniform(-1, 1, (ngrd, nz, ny, nx)).astype(np.float32) for _ in range(4)]
    pts = rng.uniform(-14, 14, (500, 3)).astype(np.float32)
    pts[:, 2] *= 0.5
    pts[-3:] = np.array([[3.0, 0.0, 0.0], [40.0, 0.0, 0.0], [0.0, 0.0, -30.0]], np.float32)
    want, wgp = O.sample_fields(pts, ngrd, mn, mx, xs, re, fields)
    assert (wgp[:, 0] > 0).sum() > 400   # most points are sampled
    dfs = [dev.from_numpy(f).cuda() for f in fields]
    ptr = (np.array([f.data_ptr() for f in dfs], np.uint64))
    import ctypes
    parr = (ctypes.c_void_p * len(dfs))(*[int(x) for x in ptr])
    dp = dev.from_numpy(np.ascontiguousarray(pts)).cuda()
    out = dev.zeros((len(pts), len(fields)), dtype=dev.float32, device="cuda")
    gp = dev.zeros((len(pts), 4), dtype=dev.int32, device="cuda")
    _lib.call("sample_fields", len(pts), dp, ngrd, mn, mx, xs, re, nx, ny, nz, len(fields), parr,
              out, gp)
    _lib.sync()
    _lib.check("sample_fields")
    assert same_bits(gp.cpu().numpy(), wgp)
    assert same_bits(out.cpu().numpy(), want)


@pytest.mark.parametrize("errlim", [1e-4, 1e-6])
def test_sparse_solver(dev, errlim):
    """The explicit-matrix CGNR entry point: same system and start as the
    oracle; the solutions agree to the solve's own accuracy (dots in double
    on the GPU vs sequential single in the reference), and the GPU meets
    the reference's stopping criterion |r|^2 <= |b|^2 npts errlim^2."""
    from g3m3_amd import _lib
    nx, ny, nz = 9, 8, 7
    npts = nx * ny * nz
    M = O.divb_matrix(nx, ny, nz)
    pos = np.array([nx * ny, nx, 1, 0, -1, -nx, -nx * ny], np.int32)
    rhs = np.random.default_rng(5).uniform(-1, 1, npts).astype(np.float32)
    so = np.zeros(npts, np.float32)
    ro = np.zeros(1, np.float32)
    O.sparse_solver(7, npts, pos, np.float32(errlim), False, 0, rhs, so, ro, M)
    dM = dev.from_numpy(np.ascontiguousarray(M)).cuda()
    drhs = dev.from_numpy(rhs).cuda()
    ds = dev.zeros(npts, dtype=dev.float32, device="cuda")
    rg = np.zeros(1, np.float32)
    _lib.call("sparse_solver", 7, npts, pos, np.float32(errlim), False, 0, drhs, ds, rg, dM)
    _lib.sync()
    _lib.check("sparse_solver")
    sg = ds.cpu().numpy()
    # the GPU's own residual, recomputed here in double
    Ad = np.zeros(npts)
    for b in range(7):
        ii = np.arange(npts) + pos[b]
        ok = (ii >= 0) & (ii < npts)
        Ad[ok] += M[ii[ok], b].astype(np.float64) * sg[ii[ok]]
    r2 = float(((Ad - rhs) ** 2).sum())
    b2 = float((rhs.astype(np.float64) ** 2).sum())
    assert r2 <= 1.5 * b2 * npts * errlim ** 2
    assert abs(rg[0] - r2) <= 1e-3 * max(r2, 1e-30) + 1e-12
    scale = float(np.abs(so).max())
    tol = max(1e-3, 20 * np.sqrt(npts) * errlim)
    assert float(np.abs(sg - so).max()) <= tol * scale


def test_divb_cg_mode(dev):
    """divB_correct in conjugate-gradient mode (mode 2): converges on the
    reference criterion in far fewer iterations than CGNR, and removes the
    divergence it solved for like the reference's solve (compared with the
    oracle's CGNR result to the solver tolerance)."""
    from g3m3_amd import _lib
    import ctypes
    w = World(17, 15, 11, 1, seed=3, lists=False)
    errlim = np.float32(1e-5)
    res = {}
    for mode in (1, 2):
        d = to_dev(dev, w)
        it = ctypes.c_longlong(0)
        r2 = ctypes.c_double(0)
        st = _lib.call("divb_correct_ex", w.nx, w.ny, w.nz, 1, 1, w.a["xspac"], errlim,
                       d["bx"], d["by"], d["bz"], mode, ctypes.c_longlong(0), it, r2)
        _lib.check("divb")
        res[mode] = (st, it.value, r2.value, {k: d[k].cpu().numpy() for k in ("bx", "by", "bz")})
    st, iters, r2, got = res[2]
    assert st == 1 and 0 < iters < res[1][1]
    # the reference criterion |A x - b|^2 <= |b|^2 npts errlim^2 on the true
    # residual (a margin for the recursive residual the CG stops on)
    bx, by, bz = (w.a[c][0].astype(np.float64) for c in ("bx", "by", "bz"))
    rhs = np.zeros_like(bx)
    rhs[1:-1, 1:-1, 1:-1] = ((bx[1:-1, 1:-1, 2:] - bx[1:-1, 1:-1, :-2])
                             + (by[1:-1, 2:, 1:-1] - by[1:-1, :-2, 1:-1])
                             + (bz[2:, 1:-1, 1:-1] - bz[:-2, 1:-1, 1:-1])) / 2.0
    npts = w.nx * w.ny * w.nz
    assert r2 <= 2.0 * float((rhs ** 2).sum()) * npts * float(errlim) ** 2
    wo = w.copy()
    pos = np.array([w.nx * w.ny, w.nx, 1, 0, -1, -w.nx, -w.nx * w.ny], np.int32)
    O.divB_correct(w.nx, w.ny, w.nz, 1, 1, 7, pos, wo.a["xspac"], errlim,
                   wo.a["bx"], wo.a["by"], wo.a["bz"])
    for k in ("bx", "by", "bz"):
        assert np.abs(got[k] - wo.a[k]).max() <= 0.02 * np.abs(w.a[k]).max()


def test_divb_refine_vs_fp64(dev):
    """Mode 3 (FP32 vectors, FP64 dots, FP64 residual refinement) against an
    all-FP64 solve: relative residual 1e-9 reached, and every field's error
    against the FP64 oracle at FP32 rounding level -- well below what the
    FP32-vector CG alone (mode 2, 1e-6) leaves."""
    pytest.importorskip("scipy.sparse")
    from g3m3_amd import _lib
    import ctypes
    w = World(21, 19, 13, 1, seed=8, lists=False)
    npts = w.nx * w.ny * w.nz
    ref = O.divB_correct_f64(w.a["bx"][0], w.a["by"][0], w.a["bz"][0], 1.0)
    err = {}
    for mode, rel in ((2, 1e-6), (3, 1e-9)):
        d = to_dev(dev, w)
        it = ctypes.c_longlong(0)
        r2 = ctypes.c_double(0)
        st = _lib.call("divb_correct_ex", w.nx, w.ny, w.nz, 1, 1, w.a["xspac"],
                       np.float32(rel / np.sqrt(npts)), d["bx"], d["by"], d["bz"], mode,
                       ctypes.c_longlong(0), it, r2)
        _lib.check("divb")
        assert st == 1, (mode, st, it.value)
        got = [d[k].cpu().numpy()[0].astype(np.float64) for k in ("bx", "by", "bz")]
        err[mode] = max(float(np.abs(g - r).max() / np.abs(r).max()) for g, r in zip(got, ref))
        print(f"divB mode {mode}: {it.value} iterations, max rel field error vs FP64 "
              f"{err[mode]:.2e}", flush=True)
    assert err[3] <= 2e-6
    assert err[3] <= err[2]


def test_find_aux_cube_host(dev):
    """The host find_aux_cube entry point against the oracle: placement
    between trajectory points, findgrid and the cube vertices."""
    from g3m3_amd import _lib
    import ctypes
    ngrd, re = 3, np.float32(0.5)
    xs = np.array([1.0, 2.0, 4.0], np.float32)
    mn = np.ascontiguousarray(np.stack([-6 * xs, -6 * xs, -3 * xs], 1).astype(np.float32))
    mx = np.ascontiguousarray(np.stack([6 * xs, 6 * xs, 3 * xs], 1).astype(np.float32))
    rng = np.random.default_rng(17)
    for q in range(60):
        x = np.append(rng.uniform(-10, 10, 3), rng.uniform(0, 10)).astype(np.float32)
        z = np.append(rng.uniform(-10, 10, 3), x[3] + rng.uniform(0.1, 5)).astype(np.float32)
        if q % 10 == 0:
            z[3] = x[3]
        ut = np.float32(x[3] + rng.uniform(-1, 6))
        res = []
        for who in ("gpu", "oracle"):
            nt, rec = np.zeros(2, np.int32), np.ones(1, np.int32)
            gp, s, cv = np.zeros(4, np.int32), np.zeros(3, np.float32), np.zeros((2, 3), np.float32)
            if who == "gpu":
                name = ctypes.create_string_buffer(b"craft02 ", 8)
                _lib.call("find_aux_cube", name, 2, 40, ngrd, mn, mx, xs, re, x, z, ut, nt, rec,
                          gp, s, cv, _lib.ByValue(ctypes.c_size_t(8)))
            else:
                O.find_aux_cube("craft02", 2, 40, ngrd, mn, mx, xs, re, x, z, ut, nt, rec, gp, s,
                                cv)
            res.append((gp, s, cv, rec))
        for u, v in zip(*res):
            assert same_bits(u, v), (q, res)


def test_sample_craft(dev):
    """Many craft placed at time ut on their trajectories and sampled in one
    launch, against the oracle's find_aux_cube placement + sampling."""
    from g3m3_amd import _lib
    import ctypes
    nx, ny, nz, ngrd = 13, 13, 7, 3
    re = np.float32(0.5)
    xs = np.array([1.0, 2.0, 4.0], np.float32)
    mn = np.ascontiguousarray(np.stack([-6 * xs, -6 * xs, -3 * xs], 1).astype(np.float32))
    mx = np.ascontiguousarray(np.stack([6 * xs, 6 * xs, 3 * xs], 1).astype(np.float32))
    rng = np.random.default_rng(23)
    fields = [rng.uniform(-1, 1, (ngrd, nz, ny, nx)).astype(np.float32) for _ in range(3)]
    nc = 400
    xc = np.concatenate([rng.uniform(-12, 12, (nc, 3)), rng.uniform(0, 5, (nc, 1))], 1)
    zc = np.concatenate([rng.uniform(-12, 12, (nc, 3)), xc[:, 3:] + rng.uniform(0.1, 5, (nc, 1))],
                        1)
    xc, zc = xc.astype(np.float32), zc.astype(np.float32)
    xc[:, 2] *= 0.5
    zc[:, 2] *= 0.5
    zc[::17, 3] = xc[::17, 3]        # same-time pairs
    ut = np.float32(4.0)
    ws, wsc, wgp = O.sample_craft(xc, zc, ut, ngrd, mn, mx, xs, re, fields)
    dfs = [dev.from_numpy(f).cuda() for f in fields]
    parr = (ctypes.c_void_p * len(dfs))(*[f.data_ptr() for f in dfs])
    dx, dz = dev.from_numpy(np.ascontiguousarray(xc)).cuda(), dev.from_numpy(
        np.ascontiguousarray(zc)).cuda()
    s = dev.zeros((nc, 3), dtype=dev.float32, device="cuda")
    out = dev.zeros((nc, len(fields)), dtype=dev.float32, device="cuda")
    gp = dev.zeros((nc, 4), dtype=dev.int32, device="cuda")
    _lib.call("sample_craft", nc, dx, dz, ut, ngrd, mn, mx, xs, re, nx, ny, nz, len(fields),
              parr, s, out, gp)
    _lib.sync()
    assert same_bits(s.cpu().numpy(), ws)
    assert same_bits(gp.cpu().numpy(), wgp)
    assert same_bits(out.cpu().numpy(), wsc)
    assert (wgp[:, 0] > 0).sum() > 200


@pytest.mark.parametrize("shift", [0.0, 0.25, -0.25])
def test_flux_counter(dev, world, shift):
    for sp in "qho":
        for c in "xyz":
            world.a[sp + "p" + c] += np.float32(shift)
    world.s["grid_spacing"] = np.float32(2.0)
    outs = ["qflux_in", "qflux_out", "hflux_in", "hflux_out", "oflux_in", "oflux_out"]
    names = ["qpx", "qpy", "qpz", "hpx", "hpy", "hpz", "opx", "opy", "opz", "nx", "ny", "nz",
             "n_grids", "box", "rmassq", "rmassh", "rmasso", "grid_spacing"] + outs
    wo = world.copy()
    for k in outs:
        wo.a[k] = np.zeros(1, np.float32)
    O.flux_counter(*wo.args(names))
    from g3m3_amd import _lib
    d = to_dev(dev, world)
    res = [np.zeros(1, np.float32) for _ in outs]
    _lib.call("flux_counter", *[d[n] if n in d else world.s[n] for n in names[:18]], *res)
    _lib.check("flux_counter")
    assert [r[0] for r in res] == [wo.a[k][0] for k in outs]


def test_divb_correct_reference_signature(dev):
    """g3m3_divb_correct with the argument list of code/sub_divB_correct.f90:
    25-27 (the reference's CGNR, single-precision scalars) on box 2 of two
    nested boxes at errlim 1e-6 against the all-FP64 exact solve of the same
    system: every corrected component within 5e-4 of the largest |b| (the
    oracle's sequential-FP32 CGNR at the same errlim lands within 9.2e-5),
    box 1 untouched; and the same solve through g3m3_divb_correct_ex reports
    convergence with residual^2 <= |b|^2 npts errlim^2, the reference's
    criterion (code/sub_sparse_solver.f90:53, 111)."""
    import ctypes
    from g3m3_amd import _lib
    w = World(17, 15, 11, 2, seed=3, lists=False)
    errlim = np.float32(1e-6)
    pos = np.array([w.nx * w.ny, w.nx, 1, 0, -1, -w.nx, -w.nx * w.ny], np.int32)
    d = to_dev(dev, w)
    print("divb_correct (reference signature) on box 2 of 17x15x11x2", flush=True)
    _lib.call("divb_correct", w.nx, w.ny, w.nz, 2, 2, 7, pos, w.a["xspac"], errlim,
              d["bx"], d["by"], d["bz"])
    _lib.sync()
    rx = float(w.a["xspac"][1])
    ref = O.divB_correct_f64(w.a["bx"][1], w.a["by"][1], w.a["bz"][1], rx)
    scale = max(float(np.abs(r).max()) for r in ref)
    for k, r in zip(("bx", "by", "bz"), ref):
        got = d[k].cpu().numpy()
        assert same_bits(got[0], w.a[k][0]), f"{k}: box 1 changed"
        assert float(np.abs(got[1].astype(np.float64) - r).max()) <= 5e-4 * scale, k
    d2 = to_dev(dev, w)
    it, r2 = ctypes.c_longlong(0), ctypes.c_double(0)
    st = _lib.call("divb_correct_ex", w.nx, w.ny, w.nz, 2, 2, w.a["xspac"], errlim,
                   d2["bx"], d2["by"], d2["bz"], 0, ctypes.c_longlong(0), it, r2)
    _lib.check("divb_correct_ex")
    X, Y, Z = (w.a[k][1] for k in ("bx", "by", "bz"))
    rhs = ((X[1:-1, 1:-1, 2:] - X[1:-1, 1:-1, :-2]) + (Y[1:-1, 2:, 1:-1] - Y[1:-1, :-2, 1:-1])
           + (Z[2:, 1:-1, 1:-1] - Z[:-2, 1:-1, 1:-1])) / np.float32(2.0 * rx)
    b2 = float((rhs.astype(np.float64) ** 2).sum())
    npts = w.nx * w.ny * w.nz
    assert st == 1 and it.value > 0
    assert r2.value <= b2 * npts * float(errlim) ** 2 * (1 + 1e-6)


def test_banded_tpose(dev):
    """g3m3_banded_tpose (code/sub_banded_tpose.f90:20-29) bit for bit against
    the oracle (itself bit-equal to the reference build), on device vectors
    of the 7-band divB matrix and on a banded matrix of random bands."""
    from g3m3_amd import _lib
    rng = np.random.default_rng(11)
    nx, ny, nz = 13, 9, 7
    npts = nx * ny * nz
    pos = np.array([nx * ny, nx, 1, 0, -1, -nx, -nx * ny], np.int32)
    for M in (O.divb_matrix(nx, ny, nz), rng.uniform(-1, 1, (npts, 7)).astype(np.float32)):
        x = rng.uniform(-1, 1, npts).astype(np.float32)
        want = np.zeros(npts, np.float32)
        O.banded_tpose(7, npts, pos, M, x, want)
        dM, dx = dev.from_numpy(M).cuda(), dev.from_numpy(x).cuda()
        dy = dev.zeros(npts, dtype=dev.float32, device="cuda")
        print(f"banded_tpose on {npts} points", flush=True)
        _lib.call("banded_tpose", 7, npts, pos, dM, dx, dy)
        _lib.sync()
        assert same_bits(dy.cpu().numpy(), want)


def test_store_array(dev, world):
    """code/sub_store_array.f90: bx(:,:,:,box) = bx0(:,:,:,box), the other
    boxes untouched."""
    names = ["bx", "bx0", "nx", "ny", "nz", "n_grids", "box"]
    wg, _ = gpu_call(dev, world, "store_array", names)
    assert_bits(wg, oracle_call(world, "store_array", names), ["bx"])


def test_fcsmooth_single_field(dev, world):
    """code/sub_fcsmooth.f90:7-242 on one field (the single-field entry point,
    beside flux_correct's all-field one)."""
    world.a["tmpx"] = np.zeros((world.nz, world.ny, world.nx), np.float32)
    world.a["tmpy"] = np.zeros_like(world.a["tmpx"])
    world.a["tmpz"] = np.zeros_like(world.a["tmpx"])
    for field, chi in (("qrho", np.float32(0.008)), ("bz", np.float32(0.004))):
        names = [field, "old" + field, "wrk" + field, "nx", "ny", "nz", "n_grids", "box",
                 chi, "tmpx", "tmpy", "tmpz"]
        wg, _ = gpu_call(dev, world, "fcsmooth", names)
        wo = oracle_call(world, "fcsmooth", names)
        assert_bits(wg, wo, [field])


@pytest.mark.parametrize("box", [1, 2])
def test_flanks_single_field(dev, nested, box):
    """code/sub_flanks.f90:5-303 on one field (g3m3_flanks), target wrk,
    coarse values from the next box's current and old arrays."""
    from g3m3_amd import _lib
    nested.s["box"] = box
    t = 0.17
    for field in ("qrho", "bx", "opresyz"):
        d = to_dev(dev, nested)
        _lib.call("flanks", d["wrk" + field], d[field], d["old" + field], nested.nx,
                  nested.ny, nested.nz, nested.ngrd, box, nested.a["t_new"],
                  nested.a["t_old"], _lib.F64(t), d["work"], *[nested.a[g] for g in GRID])
        _lib.sync()
        wo = nested.copy()
        O.flanks(*wo.args(["wrk" + field, field, "old" + field, "nx", "ny", "nz", "n_grids",
                           "box", "t_new", "t_old"]), t, wo.a["work"],
                 *[wo.a[g] for g in GRID])
        assert same_bits(d["wrk" + field].cpu().numpy(), wo.a["wrk" + field]), field
