"""Batched runs: several independent runs stepped together on several HIP
streams give exactly the states each run reaches alone (same kernels, only
the streams differ), and each run stays bit-exact with the oracle."""
import numpy as np
import pytest

from common import same_bits

from g3m3_amd.workload import FIELDS, make_workload

pytestmark = pytest.mark.gpu


def test_batch_equals_runs_alone():
    import torch
    from g3m3_amd.driver import Batch, FastRun
    specs = [dict(L=8, n_sw=3.0, v_sw=350.0, imf=(1.0, -2.0, -4.0)),
             dict(L=10, n_sw=12.0, v_sw=700.0, imf=(-3.0, 0.5, 2.0)),
             dict(L=12, n_sw=6.0, v_sw=500.0, imf=(0.0, 4.0, -1.0))]
    wls = [make_workload("single61", seed=40 + i, **sp) for i, sp in enumerate(specs)]
    steps = 3
    batch = Batch(wls, streams=2)
    batch.init_dt()
    for _ in range(steps):
        batch.full_step()
    torch.cuda.synchronize()
    for wl, rb in zip([make_workload("single61", seed=40 + i, **sp) for i, sp in enumerate(specs)],
                      batch.runs):
        ra = FastRun(wl)
        ra.init_dt()
        for _ in range(steps):
            ra.full_step()
        ha, hb = ra.state_host(), rb.state_host()
        bad = [k for k in FIELDS if not same_bits(ha[k], hb[k])]
        assert not bad, f"{wl.nx}x{wl.ny}x{wl.nz}: batched run differs on {bad}"
        assert ra.cell_updates == rb.cell_updates


def _same_run(ra, rb, label):
    for pre in ("", "old", "wrk"):
        ha, hb = ra.state_host(pre), rb.state_host(pre)
        bad = [k for k in FIELDS if not same_bits(ha[k], hb[k])]
        assert not bad, f"{label}: {pre or 'state'} arrays differ on {bad}"
    assert ra.t == rb.t and ra.cell_updates == rb.cell_updates
    assert list(ra.t_step) == list(rb.t_step) and list(ra.t_stepnew) == list(rb.t_stepnew)


@pytest.mark.parametrize("L", [10, 30])
def test_graph_steps_equal_host_steps(L):
    """Full steps replayed from captured graphs, with the time step formed
    on the device, equal the host-driven steps bit for bit -- arrays, time,
    time steps -- also when graph and host steps alternate."""
    from g3m3_amd.driver import FastRun
    wl = lambda: make_workload("single61", seed=77, L=L)
    ra, rb = FastRun(wl()), FastRun(wl())
    ra.init_dt()
    rb.init_dt()
    print(f"graph vs host steps, {ra.nx}x{ra.ny}x{ra.nz}", flush=True)
    for _ in range(5):
        ra.full_step()
    rb.graph_steps(2)
    rb.graph_finish()
    rb.full_step()
    rb.graph_steps(2)
    rb.graph_finish()
    _same_run(ra, rb, f"L={L}")


def test_batch_graph_equals_runs_alone():
    from g3m3_amd.driver import Batch, FastRun
    specs = [dict(L=8, n_sw=3.0, v_sw=350.0, imf=(1.0, -2.0, -4.0)),
             dict(L=11, n_sw=12.0, v_sw=700.0, imf=(-3.0, 0.5, 2.0)),
             dict(L=14, n_sw=6.0, v_sw=500.0, imf=(0.0, 4.0, -1.0))]
    mk = lambda: [make_workload("single61", seed=50 + i, **sp) for i, sp in enumerate(specs)]
    batch = Batch(mk(), streams=2)
    batch.init_dt()
    print("batched graph steps, 3 runs on 2 streams", flush=True)
    batch.graph_steps(4)
    for wl, rb in zip(mk(), batch.runs):
        ra = FastRun(wl)
        ra.init_dt()
        for _ in range(4):
            ra.full_step()
        _same_run(ra, rb, f"{wl.nx}x{wl.ny}x{wl.nz}")


def test_fused_batch_equals_runs_alone():
    """Runs stepped in lockstep by one multi-run launch per stage (their
    grids concatenated, a captured graph per step) reach exactly the arrays,
    time and time steps each run reaches alone -- runs of several sizes,
    one of them a step ahead (the other buffer parity) when the batch is
    made, one with a cell size that is not a power of two (its stage kernels
    divide where the others multiply by exact reciprocals, so that stage's
    launch splits by kernel), and the steps issued in two calls."""
    from g3m3_amd.driver import Batch, FastRun
    specs = [dict(L=8, n_sw=3.0, v_sw=350.0, imf=(1.0, -2.0, -4.0)),
             dict(L=9, n_sw=12.0, v_sw=700.0, imf=(-3.0, 0.5, 2.0)),
             dict(L=11, n_sw=6.0, v_sw=500.0, imf=(0.0, 4.0, -1.0)),
             dict(L=14, n_sw=1.5, v_sw=420.0, imf=(2.0, 2.0, 2.0)),
             dict(dims=(19, 17, 13), spacing=0.75, n_sw=4.0, v_sw=450.0, imf=(0.5, -1.0, 3.0))]
    mk = lambda: [make_workload("single61", seed=60 + i, **sp) for i, sp in enumerate(specs)]
    batch = Batch(mk(), streams=2)
    batch.init_dt()
    batch.runs[1].full_step()
    print("fused batch steps, 5 runs", flush=True)
    batch.fused_steps(3)
    batch.fused_steps(2)
    assert batch.fused_launches > 0
    for i, (wl, rb) in enumerate(zip(mk(), batch.runs)):
        ra = FastRun(wl)
        ra.init_dt()
        for _ in range(5 + (i == 1)):
            ra.full_step()
        _same_run(ra, rb, f"{wl.nx}x{wl.ny}x{wl.nz}")


def test_fused_batch_refuses_a_run_stepped_alone():
    """A run stepped on its own after the batch was made no longer matches
    the batch's recorded buffer rotation: the next lockstep step is refused."""
    from g3m3_amd import _lib
    from g3m3_amd.driver import Batch
    wls = [make_workload("single61", seed=70 + i, L=8) for i in range(2)]
    batch = Batch(wls, streams=1)
    batch.init_dt()
    batch.fused_steps(1)
    batch.runs[0].full_step()
    with pytest.raises(_lib.G3M3Error):
        batch.fused_steps(1)
