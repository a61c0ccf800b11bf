"""Pins the oracle's evaluator and tensor-product restatements (CPU only).

- deriv(k) against scipy.interpolate.BSpline derivatives on the same knots
  and coefficients, every degree and order (SplineEvaluator::deriv,
  spline_evaluator.hpp:395-478).
- Extrapolation rules by their definitions: Null gives 0 outside the domain
  (null_extrapolation_rule.hpp:14-26), Constant the boundary value
  (constant_extrapolation_rule.hpp:27-73), periodic splines repeat with the
  domain length (spline_evaluator.hpp:529-543).
- 2D build and evaluation (spline_builder_2d.hpp:514-651,
  spline_evaluator_2d.hpp:370-825): interpolation of the grid values, the
  separable case c = c1 (x) c2 giving f1(x) f2(y), and partial derivatives
  against central differences.
"""
import numpy as np
import pytest
import scipy.interpolate as si

import oracle as O
from cases import breaks_for


def _coef(orc, rng, periodic):
    c = rng.standard_normal((orc.size, 1))
    if periodic:
        c[orc.nbasis:] = c[: orc.degree]
    return c


@pytest.mark.parametrize("degree", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("uniform", [True, False])
def test_derivatives_vs_scipy(degree, uniform):
    if uniform:
        orc = O.Spline(degree, True, O.GREVILLE, O.GREVILLE, ncells=14)
    else:
        orc = O.Spline(degree, False, O.GREVILLE, O.GREVILLE, breaks=breaks_for("random", 14, seed=degree))
    rng = np.random.default_rng(10 + degree)
    c = _coef(orc, rng, False)
    x = rng.uniform(orc.rmin, orc.rmax, (200, 1))
    spl = si.BSpline(orc.knots(), c[:, 0], degree, extrapolate=False)
    for k in range(degree + 1):
        got = orc.eval(x, c, deriv=k)[:, 0]
        ref = spl.derivative(k)(x[:, 0]) if k else spl(x[:, 0])
        scale = max(1.0, np.abs(ref).max())
        assert np.abs(got - ref).max() <= 1e-12 * scale, k


@pytest.mark.parametrize("uniform", [True, False])
def test_extrapolation_rules(uniform):
    if uniform:
        orc = O.Spline(3, True, O.GREVILLE, O.GREVILLE, ncells=10)
    else:
        orc = O.Spline(3, False, O.GREVILLE, O.GREVILLE, breaks=breaks_for("random", 10, seed=2))
    rng = np.random.default_rng(3)
    c = _coef(orc, rng, False)
    L = orc.rmax - orc.rmin
    below = rng.uniform(orc.rmin - L, orc.rmin, (20, 1))
    above = np.nextafter(rng.uniform(orc.rmax, orc.rmax + L, (20, 1)), np.inf)
    x = np.concatenate([below - 1e-9, above])
    at = orc.eval(np.array([[orc.rmin], [orc.rmax]]), c)[:, 0]
    null = orc.eval(x, c, extrap=("null", "null"))[:, 0]
    assert np.all(null == 0.0)
    const = orc.eval(x, c, extrap=("constant", "constant"))[:, 0]
    assert np.all(const[:20] == at[0]) and np.all(const[20:] == at[1])
    mixed = orc.eval(x, c, extrap=("constant", "null"))[:, 0]
    assert np.all(mixed[:20] == at[0]) and np.all(mixed[20:] == 0.0)


@pytest.mark.parametrize("uniform", [True, False])
def test_periodic_repeats(uniform):
    if uniform:
        orc = O.Spline(3, True, O.PERIODIC, O.PERIODIC, ncells=12)
    else:
        orc = O.Spline(3, False, O.PERIODIC, O.PERIODIC, breaks=breaks_for("jitter", 12, seed=4))
    rng = np.random.default_rng(5)
    c = _coef(orc, rng, True)
    L = orc.rmax - orc.rmin
    x = rng.uniform(orc.rmin, orc.rmax, (50, 1))
    base = orc.eval(x, c)[:, 0]
    for shift in (-2, -1, 1, 3):
        got = orc.eval(x + shift * L, c)[:, 0]
        assert np.abs(got - base).max() <= 1e-12 * max(1.0, np.abs(base).max())


def _pair2d():
    o1 = O.Spline(3, False, O.GREVILLE, O.GREVILLE, breaks=breaks_for("random", 9, seed=7))
    o2 = O.Spline(3, True, O.PERIODIC, O.PERIODIC, ncells=8)
    return o1, o2


def test_2d_interpolates_grid():
    o1, o2 = _pair2d()
    rng = np.random.default_rng(6)
    vals = rng.standard_normal((2, o1.npts, o2.npts))
    coef = O.spline2d_build(o1, o2, vals)
    X, Y = np.meshgrid(o1.points(), o2.points(), indexing="ij")
    coords = np.broadcast_to(np.stack([X.ravel(), Y.ravel()], -1), (2, X.size, 2)).copy()
    got = O.spline2d_eval(o1, o2, coords, coef)
    assert np.abs(got - vals.reshape(2, -1)).max() <= 1e-13 * np.abs(vals).max()


def test_2d_separable_and_partials():
    o1, o2 = _pair2d()
    rng = np.random.default_rng(8)
    c1 = _coef(o1, rng, False)
    c2 = _coef(o2, rng, True)
    coef = (c1[:, 0][:, None] * c2[:, 0][None, :])[None]
    P = 300
    xs = rng.uniform(o1.rmin, o1.rmax, P)
    ys = rng.uniform(o2.rmin, o2.rmax, P)
    coords = np.stack([xs, ys], -1)[None]
    f1 = o1.eval(xs[:, None], c1)[:, 0]
    f2 = o2.eval(ys[:, None], c2)[:, 0]
    g1 = o1.eval(xs[:, None], c1, deriv=1)[:, 0]
    g2 = o2.eval(ys[:, None], c2, deriv=1)[:, 0]
    for (d1, d2), ref in {(0, 0): f1 * f2, (1, 0): g1 * f2, (0, 1): f1 * g2, (1, 1): g1 * g2}.items():
        got = O.spline2d_eval(o1, o2, coords, coef, d1, d2)[0]
        assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()), (d1, d2)
    # central differences of the value form, h = 1e-6
    h = 1e-6
    inner = (xs > o1.rmin + 2 * h) & (xs < o1.rmax - 2 * h)
    for k, ref in ((0, g1 * f2), (1, f1 * g2)):
        cp, cm = coords.copy(), coords.copy()
        cp[..., k] += h
        cm[..., k] -= h
        fd = (O.spline2d_eval(o1, o2, cp, coef) - O.spline2d_eval(o1, o2, cm, coef))[0] / (2 * h)
        assert np.abs(fd - ref)[inner].max() <= 1e-6 * max(1.0, np.abs(ref).max())


def _hermite_2d_data(s1, s2, P, Q, nb=1):
    """Interpolation data of P(x) Q(y) for Hermite closures in both
    dimensions: values, per-dimension boundary derivatives (row k is order
    k + degree % 2) and the corner cross derivatives."""
    x, y = s1.points(), s2.points()
    o1, o2 = s1.degree % 2, s2.degree % 2

    def d(p, k, t):
        return p.deriv(k)(t) if k else p(t)

    k1 = [k + o1 for k in range(s1.nbv_l)]
    k2 = [k + o2 for k in range(s2.nbv_l)]
    vals = np.outer(P(x), Q(y))
    dmin1 = np.array([[d(P, k, s1.rmin) * Q(yy) for yy in y] for k in k1])
    dmax1 = np.array([[d(P, k, s1.rmax) * Q(yy) for yy in y] for k in k1])
    dmin2 = np.array([[P(xx) * d(Q, k, s2.rmin) for k in k2] for xx in x])
    dmax2 = np.array([[P(xx) * d(Q, k, s2.rmax) for k in k2] for xx in x])
    mixed = {(a, b): np.array([[d(P, i, a1) * d(Q, j, b1) for j in k2] for i in k1])
             for a, a1 in (("min", s1.rmin), ("max", s1.rmax)) for b, b1 in (("min", s2.rmin), ("max", s2.rmax))}
    rep = lambda a: np.repeat(a[None], nb, axis=0)  # noqa: E731
    return (rep(vals), rep(dmin1), rep(dmax1), rep(dmin2), rep(dmax2),
            {k: rep(v) for k, v in mixed.items()})


def test_2d_hermite_build_reproduces_tensor_polynomials():
    """SplineBuilder2D with Hermite closures in both dimensions (boundary
    derivatives, and the corner cross derivatives) reproduces P(x) Q(y) of
    degree <= the spline degrees exactly: values, partial derivatives."""
    from cases import breaks_for

    s1 = O.Spline(3, False, O.HERMITE, O.HERMITE, breaks=breaks_for("random", 7, seed=2))
    s2 = O.Spline(5, True, O.HERMITE, O.HERMITE, ncells=6, rmin=-0.5, rmax=1.5)
    P = np.polynomial.Polynomial([0.4, -1.1, 0.6, 0.9])
    Q = np.polynomial.Polynomial([1.0, 0.3, -0.7, 0.2, 0.5, -0.4])
    vals, dmin1, dmax1, dmin2, dmax2, mixed = _hermite_2d_data(s1, s2, P, Q)
    coef = O.spline2d_build(s1, s2, vals, dmin1, dmax1, dmin2, dmax2, mixed["min", "min"], mixed["max", "min"],
                            mixed["min", "max"], mixed["max", "max"])
    rng = np.random.default_rng(3)
    xy = np.stack([rng.uniform(s1.rmin, s1.rmax, 300), rng.uniform(s2.rmin, s2.rmax, 300)], axis=-1)[None]
    ref = P(xy[0, :, 0]) * Q(xy[0, :, 1])
    got = O.spline2d_eval(s1, s2, xy, coef)[0]
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    dref = P.deriv()(xy[0, :, 0]) * Q.deriv()(xy[0, :, 1])
    dgot = O.spline2d_eval(s1, s2, xy, coef, 1, 1)[0]
    assert np.abs(dgot - dref).max() <= 1e-10 * np.abs(dref).max()
    # a wrong corner derivative changes the interpolant: the data is used
    bad = mixed["min", "min"].copy()
    bad[0, 0, 0] += 1.0
    coef2 = O.spline2d_build(s1, s2, vals, dmin1, dmax1, dmin2, dmax2, bad, mixed["max", "min"],
                             mixed["min", "max"], mixed["max", "max"])
    assert np.abs(coef2 - coef).max() > 1e-6


def _scipy_f(orc, c):
    """The spline of coefficients c (size() entries) as a scipy BSpline on the
    oracle's knots (values inside [rmin, rmax])."""
    return si.BSpline(orc.knots(), c, orc.degree, extrapolate=False)


@pytest.mark.parametrize("uniform", [True, False])
def test_constant_rule_eval_pos(uniform):
    """ConstantExtrapolationRule<DimI>(eval_pos) (constant_extrapolation_rule.hpp:
    39-72): points below rmin take f(eval_pos_lower), points above rmax
    f(eval_pos_upper), at interior positions, checked against scipy at those
    positions; the adjoint scatters the basis weights at eval_pos and gives
    a zero coordinate gradient there."""
    if uniform:
        orc = O.Spline(3, True, O.GREVILLE, O.GREVILLE, ncells=10)
    else:
        orc = O.Spline(3, False, O.GREVILLE, O.GREVILLE, breaks=breaks_for("random", 10, seed=2))
    rng = np.random.default_rng(4)
    c = _coef(orc, rng, False)
    L = orc.rmax - orc.rmin
    plo, phi = orc.rmin + 0.31 * L, orc.rmin + 0.77 * L
    x = np.concatenate([rng.uniform(orc.rmin - L, orc.rmin - 1e-9, (15, 1)),
                        rng.uniform(orc.rmin, orc.rmax, (10, 1)),
                        rng.uniform(orc.rmax + 1e-9, orc.rmax + L, (15, 1))])
    spl = _scipy_f(orc, c[:, 0])
    ref = spl(x[:, 0])
    ref[:15] = spl(plo)
    ref[25:] = spl(phi)
    got = orc.eval(x, c, extrap=("constant", "constant"), eval_pos=(plo, phi))[:, 0]
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
    # the bound form differs: eval_pos is used
    bound = orc.eval(x, c, extrap=("constant", "constant"))[:, 0]
    assert np.abs(bound[:15] - ref[:15]).max() > 1e-6
    w = rng.standard_normal((40, 1))
    gc, gx = orc.eval_vjp(x, w, c, extrap=("constant", "constant"), eval_pos=(plo, phi))
    dm = si.BSpline.design_matrix(np.array([plo, phi]), orc.knots(), 3).toarray()
    inside = si.BSpline.design_matrix(x[15:25, 0], orc.knots(), 3).toarray()
    gref = w[:15, 0].sum() * dm[0] + w[25:, 0].sum() * dm[1] + inside.T @ w[15:25, 0]
    assert np.abs(gc[:, 0] - gref).max() <= 1e-13 * np.abs(gref).max()
    assert np.all(gx[:15] == 0.0) and np.all(gx[25:] == 0.0)
    assert np.abs(gx[15:25, 0] - w[15:25, 0] * spl.derivative()(x[15:25, 0])).max() <= 1e-12


@pytest.mark.parametrize("dim2_periodic", [False, True])
def test_2d_constant_rule_eval_pos(dim2_periodic):
    """ConstantExtrapolationRule<DimI, DimNI>(eval_pos, min, max)
    (constant_extrapolation_rule.hpp:99-195) in SplineEvaluator2D
    (spline_evaluator_2d.hpp:679-720), on separable coefficients c1 (x) c2 so
    that f = f1(x) f2(y) with f1, f2 evaluated by scipy: a point outside in
    dimension 1 takes f1(pos) f2(y') with y' = clamp(y, min, max) (or y
    wrapped, for a periodic dimension 2); else a point outside in dimension
    2 takes f1(clamp(x, min, max)) f2(pos).  Points on all four sides and at
    the corners, where dimension 1 wins."""
    o1 = O.Spline(3, False, O.GREVILLE, O.GREVILLE, breaks=breaks_for("random", 9, seed=7))
    if dim2_periodic:
        o2 = O.Spline(3, True, O.PERIODIC, O.PERIODIC, ncells=8)
    else:
        o2 = O.Spline(3, True, O.GREVILLE, O.GREVILLE, ncells=8, rmin=-1.0, rmax=2.0)
    rng = np.random.default_rng(9)
    c1 = _coef(o1, rng, False)[:, 0]
    c2 = _coef(o2, rng, dim2_periodic)[:, 0]
    coef = np.outer(c1, c2)[None]
    f1, f2 = _scipy_f(o1, c1), _scipy_f(o2, c2)
    L1, L2 = o1.rmax - o1.rmin, o2.rmax - o2.rmin
    pos = [o1.rmin + 0.2 * L1, o1.rmin + 0.9 * L1, o2.rmin + 0.35 * L2, o2.rmin + 0.6 * L2]
    lo = [o2.rmin + 0.1 * L2, o2.rmin + 0.15 * L2, o1.rmin + 0.25 * L1, o1.rmin + 0.05 * L1]
    hi = [o2.rmin + 0.8 * L2, o2.rmin + 0.7 * L2, o1.rmin + 0.75 * L1, o1.rmin + 0.95 * L1]
    P = 400
    xs = rng.uniform(o1.rmin - 0.5 * L1, o1.rmax + 0.5 * L1, P)
    ys = rng.uniform(o2.rmin - 0.5 * L2, o2.rmax + 0.5 * L2, P)
    coords = np.stack([xs, ys], -1)[None]

    def wrap2(y):
        return y - np.floor((y - o2.rmin) / L2) * L2 if dim2_periodic else y

    ref = np.empty(P)
    for p in range(P):
        x, y = xs[p], wrap2(ys[p])
        k = 0 if x < o1.rmin else 1 if x > o1.rmax else -1
        if k < 0 and not dim2_periodic:
            k = 2 if y < o2.rmin else 3 if y > o2.rmax else -1
        if k in (0, 1):
            yy = y if dim2_periodic else min(max(y, lo[k]), hi[k])
            ref[p] = f1(pos[k]) * f2(yy)
        elif k in (2, 3):
            ref[p] = f1(min(max(x, lo[k]), hi[k])) * f2(pos[k])
        else:
            ref[p] = f1(x) * f2(y)
    got = O.spline2d_eval(o1, o2, coords, coef, rules=(1, 1, 1, 1), rule_pos=(pos, lo, hi))[0]
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    # the default positions are the bounds and the other dimension's domain
    dflt = O.spline2d_eval(o1, o2, coords, coef, rules=(1, 1, 1, 1))[0]
    rp = ([o1.rmin, o1.rmax, o2.rmin, o2.rmax], [o2.rmin, o2.rmin, o1.rmin, o1.rmin],
          [o2.rmax, o2.rmax, o1.rmax, o1.rmax])
    assert np.array_equal(dflt, O.spline2d_eval(o1, o2, coords, coef, rules=(1, 1, 1, 1), rule_pos=rp)[0])
    assert np.abs(dflt - got).max() > 1e-6
