"""The torch FP64 divB reference (tests/fp64ref.py) equals the oracle's
exact sparse solve (oracle.restate.divB_correct_f64) on small boxes."""
import numpy as np
import pytest

from fp64ref import divb_fp64


@pytest.mark.parametrize("shape,rx", [((13, 11, 9), 2.0), ((9, 17, 12), 1.0)])
def test_fp64_reference_equals_direct_solve(shape, rx):
    import torch
    from oracle import restate as O
    rng = np.random.default_rng(sum(shape))
    b = [rng.standard_normal(shape).astype(np.float32) for _ in range(3)]
    want = O.divB_correct_f64(*b, rx)
    got, it = divb_fp64(*[torch.from_numpy(x) for x in b], rx)
    assert it > 0
    for g, w in zip(got, want):
        assert float(np.abs(g.numpy() - w).max()) <= 1e-10 * float(np.abs(w).max())
