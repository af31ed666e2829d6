"""g3m3_set_speed_agrd_noalloc, the allocation-free set_speed_agrd: through
the C-ABI its results equal g3m3_set_speed_agrd's bit for bit (every field,
btot and the six CFL maxima) on two calls in a row, the second on a smaller
box, and a box beyond its 131072-cell scratch is refused with an error and
no change."""
import numpy as np
import pytest

from common import World, same_bits, set_speed_names

pytestmark = pytest.mark.gpu

SCALARS = ("pxmax", "pymax", "pzmax", "csmax", "alfmax", "fastest")


def _world(nx, ny, nz, seed):
    w = World(nx, ny, nz, 2, seed=seed)
    w.s.update(box=2, vlim=np.float32(1.2))
    r = w.rng
    # cells over the speed and sound-speed limits, so the side-list is used
    for f in ("qpx", "hpy", "opz"):
        m = r.random(w.a[f].shape) < 0.15
        w.a[f][m] *= 20
    m = r.random(w.a["epres"].shape) < 0.2
    w.a["epres"][m] *= 30
    return w


def _run(torch, w, name, check=True):
    from g3m3_amd import _lib
    names = set_speed_names("")
    host = {k: np.zeros(1, np.float32) for k in SCALARS}
    dev = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in w.a.items()
           if isinstance(v, np.ndarray) and v.ndim >= 3}
    args = []
    for n in names:
        if n in host:
            args.append(host[n])
        elif n in dev:
            args.append(dev[n])
        elif n in w.a:
            args.append(w.a[n])
        else:
            args.append(w.s[n])
    print(f"{name} on {w.nx}x{w.ny}x{w.nz}x{w.ngrd} box 2", flush=True)
    _lib.call(name, *args)
    _lib.lib().g3m3_sync()
    if check:
        _lib.check(name)
    out = {k: v.cpu().numpy() for k, v in dev.items()}
    out.update(host)
    return out


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    from g3m3_amd import _lib
    assert torch.cuda.is_available()
    _lib.lib()
    return torch


def test_noalloc_equals_allocating_entry_point(torch_dev):
    for (nx, ny, nz), seed in (((41, 37, 23), 11), ((19, 15, 11), 12)):
        w = _world(nx, ny, nz, seed)
        want = _run(torch_dev, w.copy(), "set_speed_agrd")
        got = _run(torch_dev, w.copy(), "set_speed_agrd_noalloc")
        assert set(got) == set(want)
        bad = [k for k in want if not same_bits(got[k], want[k])]
        assert not bad, f"{nx}x{ny}x{nz}: differs in {bad}"
        changed = [k for k in ("qpx", "hpy", "opz") if not same_bits(want[k], w.a[k])]
        assert changed, "no cell was flagged: the side-list went unused"
        assert want["fastest"][0] > 0


def test_noalloc_refuses_a_box_beyond_its_scratch(torch_dev):
    from g3m3_amd import _lib
    w = _world(65, 65, 33, 13)          # 139425 cells > 131072
    got = _run(torch_dev, w.copy(), "set_speed_agrd_noalloc", check=False)
    with pytest.raises(_lib.G3M3Error, match="131072"):
        _lib.check("noalloc")
    for k in ("qrho", "qpx", "epres", "btot"):
        assert same_bits(got[k], w.a[k]), k
    assert got["fastest"][0] == 0.0
