"""The differentiation-form workload: one periodic advection step on a 512 x 512
grid (uniform periodic cubic builder along x, batched over y), evaluated at the
feet x - v(y) dt with v(y) = sin(2 pi y), dt = 0.01, and the loss
L = sum w * f(feet) with w ~ N(0, 1).

The adjoint (SplineEvaluator eval_vjp -> coefficient gradient, then
SplineBuilder build_vjp, the transposed solve) gives dL/dvals, and eval_vjp's
coordinate gradient gives dL/dfeet.  Checks, for the FP64 and the
FP32-sweep + FP64-refinement forward builds:
- central differences with h = 1e-6 on 64 random entries of vals and of feet,
  relative error <= 1e-6;
- the product with the Jacobian assembled from the oracle's builds and
  evaluations of unit vectors (two full lines), relative error <= 1e-12.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 512
DT = 0.01
H = 1e-6


def _setup(cuda, precision):
    import torch
    import ddc_amd as D

    bs = D.UniformBSplines(3, 0.0, 1.0, N, periodic=True)
    builder = D.SplineBuilder(bs, precision=D.Precision(precision))
    ev = D.SplineEvaluator(bs, D.PeriodicExtrapolationRule(), D.PeriodicExtrapolationRule())
    x = builder.interpolation_points()
    y = np.arange(N) / N
    rng = np.random.default_rng(0xDDC0004)
    vals = np.sin(2 * np.pi * x)[:, None] * np.cos(2 * np.pi * y)[None, :] + 0.1 * rng.standard_normal((N, N))
    feet = x[:, None] - np.sin(2 * np.pi * y)[None, :] * DT
    w = rng.standard_normal((N, N))
    t = {k: torch.from_numpy(v).to(cuda) for k, v in (("vals", vals), ("feet", feet), ("w", w))}
    return bs, builder, ev, t, (vals, feet, w)


def _forward(bs, builder, ev, vals, feet):
    import torch

    coef = torch.empty((bs.size(), N), dtype=torch.float64, device=vals.device)
    builder(coef, vals)
    out = torch.empty_like(feet)
    ev(out, feet, coef)
    return out, coef


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32_refine"])
def test_advection_adjoint_fd(precision, cuda):
    import torch

    bs, builder, ev, t, _ = _setup(cuda, precision)
    if precision:
        assert builder.info.kernel == 2
    f, coef = _forward(bs, builder, ev, t["vals"], t["feet"])
    gx = torch.empty_like(t["feet"])
    gc = torch.empty_like(coef)
    ev.eval_vjp(t["feet"], t["w"], coef, gx, gc)
    gvals = torch.empty_like(t["vals"])
    builder.build_vjp(gc, gvals)
    gvals, gx = gvals.cpu().numpy(), gx.cpu().numpy()
    rng = np.random.default_rng(7)
    idx = [(int(i), int(j)) for i, j in zip(rng.integers(0, N, 64), rng.integers(0, N, 64))]
    fd_v, fd_x = [], []
    for i, j in idx:
        vp, vm = t["vals"].clone(), t["vals"].clone()
        vp[i, j] += H
        vm[i, j] -= H
        fp, _ = _forward(bs, builder, ev, vp, t["feet"])
        fm, _ = _forward(bs, builder, ev, vm, t["feet"])
        fd_v.append(float((t["w"] * (fp - fm)).sum()) / (2 * H))
        xp, xm = t["feet"].clone(), t["feet"].clone()
        xp[i, j] += H
        xm[i, j] -= H
        fp, _ = _forward(bs, builder, ev, t["vals"], xp)
        fm, _ = _forward(bs, builder, ev, t["vals"], xm)
        fd_x.append(float((t["w"] * (fp - fm)).sum()) / (2 * H))
    got_v = np.array([gvals[i, j] for i, j in idx])
    got_x = np.array([gx[i, j] for i, j in idx])
    assert np.abs(got_v - fd_v).max() <= 1e-6 * max(1.0, np.abs(fd_v).max())
    assert np.abs(got_x - fd_x).max() <= 1e-6 * max(1.0, np.abs(fd_x).max())


def test_advection_adjoint_jacobian(cuda):
    """dL/dvals on two lines against J^T w, J from the oracle's build and
    evaluation of every unit vector of the line."""
    import torch
    import oracle as O

    bs, builder, ev, t, (vals, feet, w) = _setup(cuda, 0)
    f, coef = _forward(bs, builder, ev, t["vals"], t["feet"])
    gx = torch.empty_like(t["feet"])
    gc = torch.empty_like(coef)
    ev.eval_vjp(t["feet"], t["w"], coef, gx, gc)
    gvals = torch.empty_like(t["vals"])
    builder.build_vjp(gc, gvals)
    gvals = gvals.cpu().numpy()
    orc = O.Spline(3, True, O.PERIODIC, O.PERIODIC, ncells=N, rmin=0.0, rmax=1.0)
    units = orc.build(np.eye(N))  # coefficients of every unit vector, one per column
    for j in (3, 300):
        J = orc.eval(np.repeat(feet[:, j][:, None], N, axis=1), units)
        ref = J.T @ w[:, j]
        assert np.abs(gvals[:, j] - ref).max() <= 1e-12 * np.abs(ref).max()
