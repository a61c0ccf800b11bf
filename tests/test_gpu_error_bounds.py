"""Interpolation error against the analytic bound, as the reference's
periodic builder tests check it (tests/splines/periodic_spline_builder.cpp:
196-207 with spline_error_bounds.hpp): for a periodic spline of degree d on
cells of width h interpolating f,

    max |f - s|   <= K_{d+1} (h / pi)^{d+1} max |f^{(d+1)}|
    max |f' - s'| <= K_d     (h / pi)^d     max |f^{(d+1)}|

(Tikhomirov 1969, Math. USSR Sb. 9 275; Volkov & Subbotin 2015, eq. 14-15),
with K_r the Favard constants, computed here from their series
K_r = (4 / pi) sum_k ((-1)^k / (2k+1))^(r+1); and the integral within the
domain length times the first bound.  f is the reference's cosine test
function, cos(2 pi (x - x0) / L + phase).  In 2D the bounds of the two
dimensions add.  These pin the GPU build and evaluation to the mathematics,
independently of the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def favard(r):
    k = np.arange(2_000_000, dtype=np.float64)
    return 4.0 / np.pi * np.sum(((-1.0) ** k / (2 * k + 1)) ** (r + 1))


def bound(h, d, norm):
    return favard(d + 1) * (h / np.pi) ** (d + 1) * norm


def test_favard_constants():
    """The series gives the known values K_1 = pi/2, K_2 = pi^2/8,
    K_3 = pi^3/24, K_4 = 5 pi^4/384."""
    for r, v in ((1, np.pi / 2), (2, np.pi ** 2 / 8), (3, np.pi ** 3 / 24), (4, 5 * np.pi ** 4 / 384)):
        assert abs(favard(r) - v) <= 1e-6 * v


@pytest.mark.parametrize("degree", range(1, 8))
@pytest.mark.parametrize("ncells", [10, 64])
def test_periodic_error_bounds(degree, ncells, cuda):
    import torch
    import ddc_amd as D

    x0, xN = -1.0, 2.0
    L = xN - x0
    w = 2 * np.pi / L
    bs = D.UniformBSplines(degree, x0, xN, ncells, periodic=True)
    builder = D.SplineBuilder(bs)
    ev = D.SplineEvaluator(bs, D.PeriodicExtrapolationRule(), D.PeriodicExtrapolationRule())
    xi = builder.interpolation_points()
    phases = np.array([0.0, 0.3, 1.1])
    B = phases.size
    f = lambda x, k=0: w ** k * np.cos(w * (x - x0) + phases + k * np.pi / 2)
    vals = torch.from_numpy(f(xi[:, None])).to(cuda)
    coef = torch.empty((bs.size(), B), dtype=torch.float64, device=cuda)
    builder(coef, vals)
    xe = np.linspace(x0, xN, 997, endpoint=False)[:, None] + np.zeros(B)
    tx = torch.from_numpy(np.ascontiguousarray(xe)).to(cuda)
    out, d1 = torch.empty_like(tx), torch.empty_like(tx)
    ev(out, tx, coef)
    ev.deriv(1, d1, tx, coef)
    integ = torch.empty(B, dtype=torch.float64, device=cuda)
    ev.integrate(integ, coef)
    h = L / ncells
    norm = w ** (degree + 1)
    err = np.abs(out.cpu().numpy() - f(xe)).max()
    err_d = np.abs(d1.cpu().numpy() - f(xe, 1)).max()
    assert err <= max(bound(h, degree, norm), 1e-14), (err, bound(h, degree, norm))
    assert err_d <= max(bound(h, degree - 1, norm), 1e-12 * w), (err_d, bound(h, degree - 1, norm))
    assert np.abs(integ.cpu().numpy()).max() <= max(L * bound(h, degree, norm), 1e-14 * L)


@pytest.mark.parametrize("d1,d2,n1,n2", [(3, 3, 16, 12), (2, 5, 20, 10), (4, 1, 9, 40)])
def test_periodic_error_bounds_2d(d1, d2, n1, n2, cuda):
    """2D (spline_builder_2d / spline_evaluator_2d tests): f = cos(w1 x) cos(w2 y)."""
    import torch
    import ddc_amd as D

    b1 = D.UniformBSplines(d1, 0.0, 1.0, n1, periodic=True)
    b2 = D.UniformBSplines(d2, 0.0, 2.0, n2, periodic=True)
    builder = D.SplineBuilder2D(D.SplineBuilder(b1), D.SplineBuilder(b2))
    ev = D.SplineEvaluator2D(D.SplineEvaluator(b1, D.PeriodicExtrapolationRule(), D.PeriodicExtrapolationRule()),
                             D.SplineEvaluator(b2, D.PeriodicExtrapolationRule(), D.PeriodicExtrapolationRule()))
    w1, w2 = 2 * np.pi, np.pi
    x = builder.builder1.interpolation_points()
    y = builder.builder2.interpolation_points()
    vals = torch.from_numpy((np.cos(w1 * x)[:, None] * np.cos(w2 * y)[None, :])[None]).to(cuda)
    coef = torch.empty((1, b1.size(), b2.size()), dtype=torch.float64, device=cuda)
    builder(coef, vals)
    rng = np.random.default_rng(2)
    pts = np.stack([rng.uniform(0, 1, 5000), rng.uniform(0, 2, 5000)], axis=-1)[None]
    out = torch.empty((1, 5000), dtype=torch.float64, device=cuda)
    ev(out, torch.from_numpy(pts).to(cuda), coef)
    err = np.abs(out.cpu().numpy()[0] - np.cos(w1 * pts[0, :, 0]) * np.cos(w2 * pts[0, :, 1])).max()
    b = bound(1.0 / n1, d1, w1 ** (d1 + 1)) + bound(2.0 / n2, d2, w2 ** (d2 + 1))
    assert err <= max(b, 1e-14), (err, b)
