"""GPU parity of the initialisation entry points -- dipole, mak_dip_grd,
set_resist and the main program's initialisation block -- against the
oracle on the same inputs.

Bar: bit-exact, except the ionospheric profile's real power
((r + r_i/2) / (1.5 r_i))**(-alpha_e) (code/3d_main.f90:1311-1312), which
the device evaluates in double precision and rounds once while the
reference calls the C library's powf: fields that depend on it agree to
4 ulp (relative 5e-7); with no ionospheric source (erho = 0) the power
does not reach any field and everything is bit-exact."""
import numpy as np
import pytest

from common import same_bits

from common import FIELDS, INIT_SCALARS, World, init_case, init_oracle
from oracle import restate as O

pytestmark = pytest.mark.gpu

F = np.float32
ROT = dict(rot_angle=0.7, xdip=0.00001, ydip=-0.3, zdip=0.2,
           sin_tilt=float(F(np.sin(0.35))), cos_tilt=float(F(np.cos(0.35))), b0=270.0)
ROT_ARGS = ("rot_angle", "xdip", "ydip", "zdip", "sin_tilt", "cos_tilt", "b0")


@pytest.fixture(scope="module")
def dev():
    import torch
    from g3m3_amd import _lib
    assert torch.cuda.is_available()
    _lib.lib()
    return torch


def test_dipole_host(dev):
    from g3m3_amd import _lib
    rng = np.random.default_rng(4)
    for p in rng.uniform(-12, 12, (300, 3)).astype(F):
        a = [np.zeros(1, F) for _ in range(3)]
        b = [np.zeros(1, F) for _ in range(3)]
        _lib.call("dipole", *a, p[0], p[1], p[2], *[F(ROT[k]) for k in ROT_ARGS])
        O.dipole(*b, p[0], p[1], p[2], **ROT)
        assert [x[0] for x in a] == [x[0] for x in b], p


@pytest.mark.parametrize("box", [1, 2])
def test_mak_dip_grd(dev, box):
    from g3m3_amd import _lib
    nx, ny, nz, ng, mb = 21, 17, 11, 2, 1
    w = World(nx, ny, nz, ng, seed=3, lists=False)
    shp = (ng, nz, ny, nx)
    ij = np.zeros((4, 3, mb), np.int32)
    ij[:, :, 0] = [[11, 9, 6], [1, 1, 1], [21, 17, 11], [2, 3, 4]]
    num = np.array([3], np.int32)
    grid = [w.a[k] for k in ("grd_xmin", "grd_xmax", "grd_ymin", "grd_ymax", "grd_zmin",
                             "grd_zmax")]
    want = [np.full(shp, 5.0, F) for _ in range(3)]
    O.mak_dip_grd(*want, nx, ny, nz, ng, mb, box, ij, num, 4, F(3.0), *grid, **ROT)
    got = [dev.full(shp, 5.0, dtype=dev.float32, device="cuda") for _ in range(3)]
    dij = dev.from_numpy(ij).cuda()
    print(f"mak_dip_grd box {box} on {nx}x{ny}x{nz}x{ng}", flush=True)
    _lib.call("mak_dip_grd", *got, nx, ny, nz, ng, mb, box, dij, num, 4, F(3.0), *grid,
              *[F(ROT[k]) for k in ROT_ARGS])
    _lib.sync()
    for g, o in zip(got, want):
        assert same_bits(g.cpu().numpy(), o)


def test_set_resist(dev):
    from g3m3_amd import _lib
    nx, ny, nz, mb = 11, 9, 8, 2
    rng = np.random.default_rng(8)

    def lst(m):
        a = np.zeros((m, 3, mb), np.int32)
        a[:, 0, :] = rng.integers(1, nx + 1, (m, mb))
        a[:, 1, :] = rng.integers(1, ny + 1, (m, mb))
        a[:, 2, :] = rng.integers(1, nz + 1, (m, mb))
        return a

    ijz, ijm, ijs = lst(6), lst(9), lst(12)
    cnt = (np.array([2, 5], np.int32), np.array([9, 4], np.int32), np.array([1, 11], np.int32))
    want = np.full((mb, nz, ny, nx), 3.0, F)
    O.set_resist(want, nx, ny, nz, mb, F(64.0), ijz, cnt[0], ijm, cnt[1], ijs, cnt[2], 12, 9, 6,
                 F(1.0))
    got = dev.full(want.shape, 3.0, dtype=dev.float32, device="cuda")
    d = [dev.from_numpy(a).cuda() for a in (ijz, ijm, ijs)]
    _lib.call("set_resist", got, nx, ny, nz, mb, F(64.0), d[0], cnt[0], d[1], cnt[1], d[2],
              cnt[2], 12, 9, 6, F(1.0))
    _lib.sync()
    assert same_bits(got.cpu().numpy(), want)


def run_init(dev, c, check=True):
    """g3m3_init_plasma on device copies of an init_case; returns host
    copies of everything it writes."""
    from g3m3_amd import _lib
    L = c["lists"]
    f = {k: dev.from_numpy(v.copy()).cuda() for k, v in c["f"].items()}
    b0 = [dev.from_numpy(v.copy()).cuda() for v in c["b0"]]
    rst = dev.from_numpy(c["rst"].copy()).cuda()
    dl = {k: dev.from_numpy(L[k].copy()).cuda() for k in ("parm_srf", "parm_mid", "parm_zero",
                                                           "ijsrf", "ijmid", "ijzero")}
    num = {k: L[k].copy() for k in ("numsrf", "nummid", "numzero")}
    print(f"init_plasma on {c['nx']}x{c['ny']}x{c['nz']}x{c['ngrd']}, erho {c['s']['erho']}",
          flush=True)
    _lib.call("init_plasma", *[f[k] for k in FIELDS], *b0, rst, c["nx"], c["ny"], c["nz"],
              c["ngrd"], c["mbndry"], c["mcap"], c["mcap"], c["mcap"], c["gmin"], c["gmax"],
              dl["parm_srf"], dl["parm_mid"], dl["parm_zero"], dl["ijsrf"], dl["ijmid"],
              dl["ijzero"], num["numsrf"], num["nummid"], num["numzero"],
              *[F(c["s"][k]) for k in INIT_SCALARS])
    _lib.lib().g3m3_sync()
    if check:
        _lib.check("init_plasma")
    out = {k: v.cpu().numpy() for k, v in f.items()}
    out.update({f"b{c}0": b.cpu().numpy() for c, b in zip("xyz", b0)})
    out["rst"] = rst.cpu().numpy()
    out.update({k: v.cpu().numpy() for k, v in dl.items()})
    out.update(num)
    return out


def oracle_out(d):
    out = dict(d["f"])
    out.update({f"b{c}0": b for c, b in zip("xyz", d["b0"])})
    out["rst"] = d["rst"]
    out.update(d["lists"])
    return out


@pytest.mark.parametrize("shape", [(21, 19, 13, 2), (35, 33, 17, 3)])
def test_init_plasma_bitexact(dev, shape):
    c = init_case(*shape, erho=0.0, mcap=2000)
    got = run_init(dev, c)
    want = oracle_out(init_oracle(c)[0])
    bad = [k for k in want if not same_bits(got[k], want[k])]
    assert not bad, f"differs from the oracle: {bad}"
    assert got["numsrf"][0] > 0 and got["nummid"][0] > 0 and got["numzero"][0] > 0


def test_init_plasma_ionosphere(dev):
    """erho > 0: the ionospheric profile's power reaches the densities and
    pressures (4-ulp bar); lists, counts, b0 and rst stay bit-exact."""
    c = init_case(erho=50.0)
    got = run_init(dev, c)
    want = oracle_out(init_oracle(c)[0])
    exact = ["bx0", "by0", "bz0", "rst", "ijsrf", "ijmid", "ijzero", "numsrf", "nummid",
             "numzero"]
    for k in exact:
        assert same_bits(got[k], want[k]), k
    for k in want:
        if k in exact:
            continue
        a, b = got[k].astype(np.float64), want[k].astype(np.float64)
        assert np.all(np.abs(a - b) <= 5e-7 * np.abs(b) + 1e-30), k
    assert (want["qrho"] > 0.011).any()   # the ionosphere is above the floor somewhere


def test_init_plasma_list_capacity(dev):
    """Lists longer than their capacity are cut there and the error latch
    is set; the counts report the capacity."""
    from g3m3_amd import _lib
    c = init_case(mcap=10)
    _lib.check("before")
    got = run_init(dev, c, check=False)
    with pytest.raises(_lib.G3M3Error):
        _lib.check("init_plasma")
    assert got["numsrf"][0] == 10
    want = init_oracle(init_case(mcap=2000))[0]["lists"]
    assert same_bits(got["ijsrf"][:10], want["ijsrf"][:10])


@pytest.mark.parametrize("tilt", [0.0, 0.2])
def test_torus_inject(dev, tilt):
    """The torus source against the oracle: untouched cells bit-exact, the
    injected cells within 4 ulp (the Gaussians' exp: double on the device,
    the C library's expf in the oracle), totals to 1e-6."""
    from g3m3_amd import _lib
    from common import TORUS_SCALARS, torus_case, torus_oracle
    c = torus_case(tilt=tilt)
    d, want_tot = torus_oracle(c)
    f = {k: dev.from_numpy(v.copy()).cuda() for k, v in c["f"].items()}
    b0 = [dev.from_numpy(v.copy()).cuda() for v in c["b0"]]
    tot = [np.zeros(1, F) for _ in range(3)]
    print(f"torus on {c['nx']}x{c['ny']}x{c['nz']}, tilt {tilt}", flush=True)
    _lib.call("torus_inject", *[f[k] for k in FIELDS[:31]], f["bx"], f["by"], f["bz"], *b0,
              c["nx"], c["ny"], c["nz"], c["ngrd"], c["gmin"], c["gmax"],
              *[F(c["s"][k]) for k in TORUS_SCALARS], *tot)
    _lib.check("torus_inject")
    for k in FIELDS:
        got, want = f[k].cpu().numpy(), d["f"][k]
        same = c["f"][k] == want            # cells the source did not touch
        assert same_bits(got[same], want[same]), k
        a, b = got.astype(np.float64), want.astype(np.float64)
        assert np.all(np.abs(a - b) <= 5e-7 * np.abs(b) + 1e-30), k
    for g, w in zip(tot, want_tot):
        assert abs(float(g[0]) - float(w)) <= 1e-6 * max(abs(float(w)), 1e-30)
    if tilt == 0.0:
        assert want_tot[1] > 0
