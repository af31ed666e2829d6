"""The main loop stops on a time step below t_step_min = 1e-6
(code/3d_main.f90:116, 1969-1973); a step that is not a finite number stops
it too.  CPU checks of the oracle's step driver; the GPU drivers have the
same checks in test_gpu_steps.py."""
import numpy as np
import pytest

from g3m3_amd.workload import make_workload
from oracle.boxstep import OracleRun


@pytest.mark.parametrize("dt", [1e-7, 0.0, float("nan"), float("inf")])
def test_oracle_refuses_a_collapsed_step(dt):
    o = OracleRun(make_workload("tiny"), "restate")
    o.init_dt()
    before = {k: v.copy() for k, v in o.state().items()}
    with pytest.raises(FloatingPointError, match="t_step_min"):
        o.full_step(dt_override=[dt] * o.ngrd)
    assert o.box_steps == 0
    assert all(np.array_equal(before[k], v) for k, v in o.state().items())


def test_oracle_takes_a_step_at_t_step_min():
    o = OracleRun(make_workload("tiny"), "restate")
    o.init_dt()
    o.full_step(dt_override=[1e-6] * o.ngrd)
    assert o.box_steps > 0
