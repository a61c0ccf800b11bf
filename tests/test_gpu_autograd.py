"""torch.autograd wrappers (ddc_amd.autograd) checked by
torch.autograd.gradcheck: central differences of the HIP forward against the
HIP vjp kernels, in FP64, on a reduced shape of the advection workload
(uniform periodic cubic along x, batched over y, feet x_i - v(y) dt), for
the values and the feet together; and a Hermite case for the boundary
derivative inputs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_gradcheck_advection_values_and_feet(cuda):
    import torch
    import ddc_amd as D
    from ddc_amd.autograd import interpolate

    nx, ny = 32, 4
    bs = D.UniformBSplines(3, 0.0, 1.0, nx, periodic=True)
    builder = D.SplineBuilder(bs)
    ev = D.SplineEvaluator(bs)
    x = builder.interpolation_points()
    y = np.arange(ny) / ny
    rng = np.random.default_rng(0)
    vals = np.sin(2 * np.pi * x[:, None] * rng.integers(1, 4, ny)[None, :]) + 0.1 * rng.standard_normal((nx, ny))
    feet = x[:, None] - np.sin(2 * np.pi * y)[None, :] * 0.01
    tv = torch.tensor(vals, device=cuda, requires_grad=True)
    tf = torch.tensor(feet, device=cuda, requires_grad=True)
    assert torch.autograd.gradcheck(lambda v, f: interpolate(builder, ev, v, f), (tv, tf), eps=1e-6, atol=1e-7,
                                    rtol=1e-5, nondet_tol=1e-12)


def test_gradcheck_hermite_derivs(cuda):
    """Values and both Hermite boundary-derivative inputs, non-uniform quintic,
    interpolation-contiguous layout."""
    import torch
    import ddc_amd as D
    from ddc_amd.autograd import interpolate
    from cases import breaks_for

    br = breaks_for("random", 12, seed=3)
    bs = D.NonUniformBSplines(5, br)
    H = D.SplineBuilderClosure.HERMITE
    builder = D.SplineBuilder(bs, H, H)
    ev = D.SplineEvaluator(bs, D.NullExtrapolationRule(), D.NullExtrapolationRule())
    nb, npts = 3, 40
    rng = np.random.default_rng(1)
    vals = torch.tensor(rng.standard_normal((nb, builder.npoints())), device=cuda, requires_grad=True)
    dl = torch.tensor(rng.standard_normal((nb, 2)), device=cuda, requires_grad=True)
    dr = torch.tensor(rng.standard_normal((nb, 2)), device=cuda, requires_grad=True)
    pts = torch.tensor(rng.uniform(0.02, 0.98, (nb, npts)), device=cuda, requires_grad=True)
    assert torch.autograd.gradcheck(lambda v, a, b, p: interpolate(builder, ev, v, p, a, b, axis=1),
                                    (vals, dl, dr, pts), eps=1e-6, atol=1e-7, rtol=1e-5,
                                    nondet_tol=1e-12)
