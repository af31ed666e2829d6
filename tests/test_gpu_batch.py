"""Batched runs: several independent runs stepped together on several HIP
streams give exactly the states each run reaches alone (same kernels, only
the streams differ), and each run stays bit-exact with the oracle."""
import numpy as np
import pytest

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
        bad = [k for k in FIELDS if not np.array_equal(ha[k], hb[k])]
        assert not bad, f"{wl.nx}x{wl.ny}x{wl.nz}: batched run differs on {bad}"
        assert ra.cell_updates == rb.cell_updates
