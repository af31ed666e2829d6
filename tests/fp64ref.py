"""An all-FP64 divB clean-up in torch, for grids too large for the oracle's
sparse solves (the 241x241x121 precision configuration).

Same system as code/sub_divB_correct.f90:85-172: the right side is the
central-difference divergence / (2 rx) on the interior, the matrix the
7-point Laplacian (-6 on the diagonal, +1 for each neighbour inside the
box), phi = rx^2 x, and B -= grad phi on the interior.  Solved by
conjugate gradients on -A (symmetric positive definite) in float64 to a
relative residual ``rtol``.  Checked against oracle.restate.divB_correct_f64
(an exact sparse solve) in tests/test_fp64ref.py."""
from __future__ import annotations


def laplacian(v):
    import torch
    out = -6.0 * v
    out[:, :, 1:] += v[:, :, :-1]
    out[:, :, :-1] += v[:, :, 1:]
    out[:, 1:, :] += v[:, :-1, :]
    out[:, :-1, :] += v[:, 1:, :]
    out[1:, :, :] += v[:-1, :, :]
    out[:-1, :, :] += v[1:, :, :]
    return out


def divb_fp64(bx, by, bz, rx: float, rtol: float = 1e-13, maxiter: int = 100000):
    """bx, by, bz: (nz, ny, nx) tensors (any float dtype, any device).
    Returns the corrected fields (float64) and the CG iterations taken."""
    import torch
    X, Y, Z = (t.to(torch.float64) for t in (bx, by, bz))
    C = (slice(1, -1),) * 3
    rhs = torch.zeros_like(X)
    rhs[C] = ((X[1:-1, 1:-1, 2:] - X[1:-1, 1:-1, :-2]) + (Y[1:-1, 2:, 1:-1] - Y[1:-1, :-2, 1:-1])
              + (Z[2:, 1:-1, 1:-1] - Z[:-2, 1:-1, 1:-1])) / (2.0 * rx)
    b = -rhs
    x = torch.zeros_like(b)
    r = b.clone()
    p = r.clone()
    rr = float((r * r).sum())
    stop = (rtol ** 2) * float((b * b).sum())
    it = 0
    while rr > stop and it < maxiter:
        q = -laplacian(p)
        alpha = rr / float((p * q).sum())
        x += alpha * p
        r -= alpha * q
        rr_new = float((r * r).sum())
        p = r + (rr_new / rr) * p
        rr = rr_new
        it += 1
    if rr > stop:
        raise RuntimeError("FP64 CG did not converge")
    phi = x * (rx * rx)
    out = [X.clone(), Y.clone(), Z.clone()]
    out[0][C] -= (phi[1:-1, 1:-1, 2:] - phi[1:-1, 1:-1, :-2]) / (2.0 * rx)
    out[1][C] -= (phi[1:-1, 2:, 1:-1] - phi[1:-1, :-2, 1:-1]) / (2.0 * rx)
    out[2][C] -= (phi[2:, 1:-1, 1:-1] - phi[:-2, 1:-1, 1:-1]) / (2.0 * rx)
    return out, it
