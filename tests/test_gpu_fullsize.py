"""Full-size checks at the FP64 workload's shape: 1024 cells x 65536 lines
(512 MiB of values), both layouts, for the uniform periodic cubic and quintic
and the non-uniform cubic Greville bases.  The whole batch goes through the
same calls as the bench (so the 32-bit buffer offsets of the streaming
evaluator and the tiled kernels run at their largest), and is checked by
properties that do not need the oracle on every line:
  - build then eval at the interpolation points returns the values (<= 1e-13);
  - deriv(1) matches an exact-for-polynomials central-difference stencil of
    eval wherever the stencil stays within one cell;
and, on 256 lines spread over the batch (both ends included), the
coefficients, values and deriv(1) against the oracle (<= 1e-13)."""
import numpy as np
import pytest

from cases import Case, G, P
from conftest import rel_err

pytestmark = pytest.mark.gpu

NC, B = 1024, 65536
CASES = [Case(3, True, P, ncells=NC), Case(5, True, P, ncells=NC), Case(3, False, G, ncells=NC, mesh="random")]


@pytest.mark.parametrize("case", CASES, ids=lambda c: c.id)
@pytest.mark.parametrize("layout", ["batch_contiguous", "interp_contiguous"])
def test_full_size_build_eval(case, layout, cuda):
    import torch
    import ddc_amd as D

    torch.cuda.empty_cache()
    orc = case.oracle()
    bs, builder = case.product()
    ev = D.SplineEvaluator(bs, D.ConstantExtrapolationRule(), D.ConstantExtrapolationRule())
    bc = layout == "batch_contiguous"
    ax = 0 if bc else 1
    m, nc = builder.npoints(), bs.size()
    gen = torch.Generator(device=cuda)
    gen.manual_seed(0xDDC0002)
    xg = torch.from_numpy(builder.interpolation_points()).to(cuda)
    k = torch.randint(1, 9, (B,), device=cuda, generator=gen).to(torch.float64)
    vals = torch.sin(2 * np.pi * xg[:, None] * k[None, :]) + 0.1 * torch.randn((m, B), device=cuda,
                                                                               dtype=torch.float64, generator=gen)
    if not bc:
        vals = vals.T.contiguous()
    coef = torch.empty((nc, B) if bc else (B, nc), device=cuda, dtype=torch.float64)
    builder(coef, vals, axis=ax)
    # 1. values back at the interpolation points, every line
    pts = (xg[:, None].expand(m, B) if bc else xg[None, :].expand(B, m)).contiguous()
    out = torch.empty_like(pts)
    ev(out, pts, coef, axis=ax)
    back = ((out - vals).abs().amax() / vals.abs().amax()).item()
    assert back <= 1e-13, back
    del pts
    # 2. deriv(1) against central differences, every line.  The points are
    # multiples of h = 2^-30, so x + k h is exact, and the 6th-order stencil
    # (45 (f1 - f-1) - 9 (f2 - f-2) + (f3 - f-3)) / (60 h) is exact for
    # polynomials of degree <= 6: on points whose stencil stays in one cell
    # it differs from f' by rounding only (~ eps max|f| / h).
    lo, hi = bs.rmin(), bs.rmax()
    h = 2.0 ** -30
    d = case.degree
    br = torch.from_numpy(np.ascontiguousarray(bs.knots()[d: d + bs.ncells() + 1])).to(cuda)
    x = torch.floor((lo + (hi - lo) * torch.rand(vals.shape, device=cuda, dtype=torch.float64, generator=gen)) / h) * h
    d1 = torch.empty_like(x)
    ev.deriv(1, d1, x, coef, axis=ax)
    fd = torch.zeros_like(x)
    f = torch.empty_like(x)
    for k, c in ((1, 45.0), (2, -9.0), (3, 1.0)):
        ev(f, x + k * h, coef, axis=ax)
        fd += c * f
        ev(f, x - k * h, coef, axis=ax)
        fd -= c * f
    fd /= 60 * h
    one_cell = (torch.searchsorted(br, (x - 3 * h).contiguous(), right=True)
                == torch.searchsorted(br, (x + 3 * h).contiguous(), right=True)) & (x - 3 * h > lo) & (x + 3 * h < hi)
    frac_used = one_cell.double().mean().item()
    fd_err = ((d1 - fd).abs() * one_cell).amax().item() / d1.abs().amax().item()
    assert frac_used > 0.9, frac_used
    assert fd_err <= 5e-8, fd_err
    del f, fd, one_cell
    # 3. sampled lines against the oracle
    rng = np.random.default_rng(1)
    lines = np.unique(np.concatenate([np.arange(64), np.arange(B - 64, B), rng.choice(B, 128, replace=False)]))
    tl = torch.from_numpy(lines).to(cuda)
    take = (lambda t: t[:, tl]) if bc else (lambda t: t[tl, :].T)
    v_s, c_s, x_s = take(vals).cpu().numpy(), take(coef).cpu().numpy(), take(x).cpu().numpy()
    ref_c = orc.build(np.ascontiguousarray(v_s))
    assert rel_err(c_s, ref_c) <= 1e-13
    ev(out, x, coef, axis=ax)
    assert rel_err(take(out).cpu().numpy(), orc.eval(x_s, ref_c, extrap=("constant", "constant"))) <= 1e-13
    assert rel_err(take(d1).cpu().numpy(), orc.eval(x_s, ref_c, deriv=1)) <= 1e-13
    print(f"{case.id} {layout}: back {back:.2e}, fd {fd_err:.2e} on {frac_used:.3f} of the points")
