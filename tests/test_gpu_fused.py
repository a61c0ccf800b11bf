"""GPU parity of the fused build + evaluation entry point (ddcx_build_eval):
coefficients and values against the oracle's build and evaluation of the
same inputs, <= 1e-13 relative (max norm), with and without the coefficients
leaving the chip; the two-call path for the cases the fused kernel does not
take; and the error when it would need a coefficient array it was not given.
"""
import numpy as np
import pytest

from cases import Case, G, H, P, STRETCHED_CASES, derivs, values
from conftest import rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-13

FUSED_CASES = [
    Case(3, True, P, ncells=1024),
    Case(2, True, P, ncells=700),
    Case(5, True, P, ncells=1024),
    Case(4, True, P, ncells=500),
    Case(3, False, G, ncells=1024, mesh="random"),
    Case(3, False, H, ncells=900, mesh="random"),
    Case(5, False, G, ncells=600, mesh="random", seed=3),
    Case(4, True, G, ncells=800),
    Case(2, False, P, ncells=400, mesh="jitter"),
    STRETCHED_CASES[0],
]


def _inputs(case, orc, nbatch, npts, seed):
    rng = np.random.default_rng(seed)
    f = values(orc.npts, nbatch, seed, x=orc.points())
    dl = derivs(orc.nbv_l, nbatch, seed) if case.bcl == H else None
    dr = derivs(orc.nbv_r, nbatch, seed + 7) if case.bcu == H else None
    L = orc.rmax - orc.rmin
    x = rng.uniform(orc.rmin - 0.05 * L, orc.rmax + 0.05 * L, (npts, nbatch))
    return f, dl, dr, x


@pytest.mark.parametrize("case", FUSED_CASES, ids=lambda c: c.id)
@pytest.mark.parametrize("store", [True, False], ids=["coef", "on_chip"])
def test_build_eval_fused(case, store, cuda):
    import torch
    import ddc_amd as D

    orc = case.oracle()
    bs, builder = case.product()
    ev = D.SplineEvaluator(bs, D.ConstantExtrapolationRule(), D.NullExtrapolationRule())
    nbatch, npts = 37, 300
    f, dl, dr, x = _inputs(case, orc, nbatch, npts, 5)
    coef_ref = orc.build(f, dl, dr)
    out_ref = orc.eval(x, coef_ref, extrap=("constant", "null"))
    t = lambda a: None if a is None else torch.from_numpy(a).to(cuda)  # noqa: E731
    coef = torch.full((orc.size, nbatch), np.nan, dtype=torch.float64, device=cuda) if store else None
    out = torch.full((npts, nbatch), np.nan, dtype=torch.float64, device=cuda)
    D.build_eval(builder, ev, out, t(x), t(f), coef, t(dl), t(dr))
    assert rel_err(out.cpu().numpy(), out_ref) <= TOL
    if store:
        assert rel_err(coef.cpu().numpy(), coef_ref) <= TOL


def test_build_eval_uncovered_case(cuda):
    """Degree 5 on 2047-row main blocks: the row table and the coefficient
    tile do not fit in LDS together, so the two calls run (coef required)."""
    import torch
    import ddc_amd as D

    case = STRETCHED_CASES[1]
    orc = case.oracle()
    bs, builder = case.product()
    ev = D.SplineEvaluator(bs, D.ConstantExtrapolationRule(), D.ConstantExtrapolationRule())
    f, dl, dr, x = _inputs(case, orc, 9, 100, 8)
    t = lambda a: torch.from_numpy(a).to(cuda)  # noqa: E731
    out = torch.empty((100, 9), dtype=torch.float64, device=cuda)
    with pytest.raises(D.DDCXError):
        D.build_eval(builder, ev, out, t(x), t(f), None, t(dl), t(dr))
    coef = torch.empty((orc.size, 9), dtype=torch.float64, device=cuda)
    D.build_eval(builder, ev, out, t(x), t(f), coef, t(dl), t(dr))
    coef_ref = orc.build(f, dl, dr)
    assert rel_err(coef.cpu().numpy(), coef_ref) <= TOL
    assert rel_err(out.cpu().numpy(), orc.eval(x, coef_ref, extrap=("constant", "constant"))) <= TOL


def test_build_eval_two_call_path(cuda):
    """Interpolation-contiguous arrays take the two calls (coef required)."""
    import torch
    import ddc_amd as D

    case = Case(3, True, P, ncells=300)
    orc = case.oracle()
    bs, builder = case.product()
    ev = D.SplineEvaluator(bs, D.PeriodicExtrapolationRule(), D.PeriodicExtrapolationRule())
    f, _, _, x = _inputs(case, orc, 20, 100, 6)
    coef_ref = orc.build(f)
    out_ref = orc.eval(x, coef_ref)
    tf = torch.from_numpy(np.ascontiguousarray(f.T)).to(cuda)
    tx = torch.from_numpy(np.ascontiguousarray(x.T)).to(cuda)
    out = torch.empty_like(tx)
    with pytest.raises(D.DDCXError):
        D.build_eval(builder, ev, out, tx, tf, None, axis=1)
    coef = torch.empty((20, orc.size), dtype=torch.float64, device=cuda)
    D.build_eval(builder, ev, out, tx, tf, coef, axis=1)
    assert rel_err(out.cpu().numpy().T, out_ref) <= TOL
    assert rel_err(coef.cpu().numpy().T, coef_ref) <= TOL


@pytest.mark.parametrize("case", [FUSED_CASES[i] for i in (0, 2, 4, 5)], ids=lambda c: c.id)
@pytest.mark.parametrize("store", [True, False], ids=["coef", "on_chip"])
def test_advect_step(case, store, cuda):
    """ddcx_advect (characteristics_advection.cpp:237-274): the spline of vals
    evaluated at the feet x_i - shift[l] of the interpolation points, with the
    feet computed in the fused kernel.  Against the oracle's build and
    evaluation at the same feet (<= 1e-13), and bit for bit against
    build_eval given those feet as points."""
    import torch
    import ddc_amd as D

    orc = case.oracle()
    bs, builder = case.product()
    ev = D.SplineEvaluator(bs, D.ConstantExtrapolationRule(), D.NullExtrapolationRule())
    nbatch = 41
    f, dl, dr, _ = _inputs(case, orc, nbatch, 1, 9)
    rng = np.random.default_rng(10)
    L = orc.rmax - orc.rmin
    shift = rng.uniform(-0.03 * L, 0.03 * L, nbatch)
    feet = builder.interpolation_points()[:, None] - shift[None, :]
    coef_ref = orc.build(f, dl, dr)
    out_ref = orc.eval(feet, coef_ref, extrap=("constant", "null"))
    t = lambda a: None if a is None else torch.from_numpy(a).to(cuda)  # noqa: E731
    coef = torch.full((orc.size, nbatch), np.nan, dtype=torch.float64, device=cuda) if store else None
    out = torch.full((orc.npts, nbatch), np.nan, dtype=torch.float64, device=cuda)
    D.advect(builder, ev, out, t(f), t(shift), coef, t(dl), t(dr))
    assert rel_err(out.cpu().numpy(), out_ref) <= TOL
    if store:
        assert rel_err(coef.cpu().numpy(), coef_ref) <= TOL
    out2 = torch.empty_like(out)
    D.build_eval(builder, ev, out2, t(feet), t(f), None, t(dl), t(dr))
    assert torch.equal(out, out2)


def test_advect_config4_step(cuda):
    """The 512 x 512 periodic advection step: feet x - sin(2 pi y) dt with
    dt = 0.01, against a separate build and evaluation at the same feet."""
    import torch
    import ddc_amd as D

    n = 512
    bs = D.UniformBSplines(3, 0.0, 1.0, n, periodic=True)
    builder = D.SplineBuilder(bs)
    ev = D.SplineEvaluator(bs, D.PeriodicExtrapolationRule(), D.PeriodicExtrapolationRule())
    x = torch.from_numpy(builder.interpolation_points()).to(cuda)
    y = torch.arange(n, dtype=torch.float64, device=cuda) / n
    vals = torch.sin(2 * np.pi * x)[:, None] * torch.cos(2 * np.pi * y)[None, :]
    shift = (torch.sin(2 * np.pi * y) * 0.01).contiguous()
    out = torch.empty_like(vals)
    D.advect(builder, ev, out, vals, shift)
    coef = torch.empty((bs.size(), n), dtype=torch.float64, device=cuda)
    builder(coef, vals)
    ref = torch.empty_like(vals)
    ev(ref, (x[:, None] - shift[None, :]).contiguous(), coef)
    assert rel_err(out.cpu().numpy(), ref.cpu().numpy()) <= TOL


def test_advect_uncovered_case(cuda):
    """Interpolation-contiguous arrays are not covered by the fused kernel:
    the advection step reports it instead of running another path."""
    import torch
    import ddc_amd as D

    case = Case(3, True, P, ncells=300)
    orc = case.oracle()
    bs, builder = case.product()
    ev = D.SplineEvaluator(bs, D.PeriodicExtrapolationRule(), D.PeriodicExtrapolationRule())
    f, _, _, _ = _inputs(case, orc, 20, 1, 6)
    tf = torch.from_numpy(np.ascontiguousarray(f.T)).to(cuda)
    out = torch.empty_like(tf)
    with pytest.raises(D.DDCXError):
        D.advect(builder, ev, out, tf, torch.zeros(20, dtype=torch.float64, device=cuda), axis=1)
