"""Pins the oracle's SplineBuilder restatement (CPU only).

- Interpolants vs scipy.interpolate.make_interp_spline / BSpline (periodic,
  Greville knots, Hermite end derivatives): the same spline space and
  conditions solved by an independent code.
- Polynomial reproduction with the reference's tolerances
  (tests/splines/non_periodic_spline_builder.cpp:297-301), seeded instead of
  std::random_device.
- Quadrature coefficients integrate polynomials exactly
  (SplineBuilder::quadrature_coefficients, spline_builder.hpp:1111-1234).
- Grid validation errors (tests/splines/spline_builder.cpp:36-135) and the
  Greville ordering of periodic_spline_builder_ordered_points.cpp:41-67.
- Batch-layout permutations give identical coefficients.
"""
import numpy as np
import pytest
import scipy.interpolate as si

import oracle as O
from cases import breaks_for


def _bspline(orc, coef):
    return si.BSpline(orc.knots(), coef, orc.degree, extrapolate=False)


@pytest.mark.parametrize("degree", [1, 2, 3, 4, 5])
def test_periodic_uniform_vs_scipy(degree):
    orc = O.Spline(degree, True, O.PERIODIC, O.PERIODIC, ncells=24)
    x = orc.points()
    f = np.sin(2 * np.pi * x) + 0.3 * np.cos(6 * np.pi * x)
    c = orc.build(f[:, None])[:, 0]
    # the periodic interpolant is unique: compare values with scipy's
    xp = np.append(x, x[0] + 1.0) if x[0] == 0.0 else None
    xe = np.random.default_rng(0).uniform(orc.rmin, orc.rmax, 300)
    mine = orc.eval(xe[:, None], c[:, None])[:, 0]
    if degree % 2 == 1:
        ref = si.make_interp_spline(xp, np.append(f, f[0]), k=degree, bc_type="periodic")(xe)
        assert np.abs(mine - ref).max() <= 1e-12
    # interpolation conditions hold for every degree
    assert np.abs(orc.eval(x[:, None], c[:, None])[:, 0] - f).max() <= 1e-13


@pytest.mark.parametrize("degree", [2, 3, 4, 5])
@pytest.mark.parametrize("mesh", ["uniform", "random", "geometric"])
def test_greville_vs_scipy(degree, mesh):
    br = breaks_for(mesh, 30, seed=degree, ratio=1.1)
    orc = O.Spline(degree, False, O.GREVILLE, O.GREVILLE, breaks=br)
    x = orc.points()
    f = np.cos(3 * x) + x ** 2
    c = orc.build(f[:, None])[:, 0]
    ref = si.make_interp_spline(x, f, k=degree, t=orc.knots())
    assert np.abs(c - ref.c).max() <= 1e-12 * np.abs(ref.c).max()


@pytest.mark.parametrize("degree", [3, 5])
@pytest.mark.parametrize("mesh", ["uniform", "random"])
def test_hermite_vs_scipy(degree, mesh):
    br = breaks_for(mesh, 25, seed=1)
    orc = O.Spline(degree, False, O.HERMITE, O.HERMITE, breaks=br)
    x = orc.points()
    f = np.exp(x) * np.sin(2 * x)
    nbe = degree // 2
    g = lambda z, k: np.real(np.exp(z) * np.exp(2j * z) * (1 + 2j) ** k) if False else None  # noqa: E731
    # derivatives of exp(x) sin(2x): imaginary part of exp((1+2i)x) (1+2i)^k
    der = lambda z, k: np.imag(np.exp((1 + 2j) * z) * (1 + 2j) ** k)  # noqa: E731
    dl = np.array([[der(0.0, k)] for k in range(1, nbe + 1)])
    dr = np.array([[der(1.0, k)] for k in range(1, nbe + 1)])
    c = orc.build(f[:, None], dl, dr)[:, 0]
    spl = _bspline(orc, c)
    assert np.abs(spl(x) - f).max() <= 1e-12
    for k in range(1, nbe + 1):
        assert abs(spl(0.0, nu=k) - dl[k - 1, 0]) <= 1e-9 * max(1, abs(dl[k - 1, 0]))
        assert abs(spl(1.0, nu=k) - dr[k - 1, 0]) <= 1e-9 * max(1, abs(dr[k - 1, 0]))
    # and the same coefficients as a dense solve of the collocation system
    # assembled from scipy's B-spline derivatives (independent of the oracle)
    t = orc.knots()
    n = orc.nbasis
    rows, rhs = [], []
    for k in range(1, nbe + 1):
        rows.append([si.BSpline(t, np.eye(n)[j], degree)(0.0, nu=k) for j in range(n)])
        rhs.append(dl[k - 1, 0])
    rows += list(si.BSpline.design_matrix(x, t, degree).toarray())
    rhs += list(f)
    for k in range(1, nbe + 1):
        rows.append([si.BSpline(t, np.eye(n)[j], degree)(1.0, nu=k) for j in range(n)])
        rhs.append(dr[k - 1, 0])
    ref = np.linalg.solve(np.array(rows), np.array(rhs))
    assert np.abs(c - ref).max() <= 1e-11 * np.abs(ref).max()


@pytest.mark.parametrize("degree", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("uniform", [True, False])
@pytest.mark.parametrize("bc", ["greville", "hermite", "homogeneous"])
def test_polynomial_reproduction(degree, uniform, bc):
    """non_periodic_spline_builder.cpp:297-301 tolerances: values 1e-14, derivatives 1e-11, integral 1e-14."""
    if bc != "greville" and degree < 2:
        pytest.skip("no boundary equations for degree 1")
    rng = np.random.default_rng(100 * degree + uniform)
    coeffs = rng.uniform(-1.0, 1.0, degree + 1)  # polynomial of exact degree
    p = np.polynomial.Polynomial(coeffs)
    code = {"greville": O.GREVILLE, "hermite": O.HERMITE, "homogeneous": O.HOMOGENEOUS_HERMITE}[bc]
    br = None if uniform else breaks_for("jitter", 10, seed=degree)
    if bc == "homogeneous":
        # homogeneous Hermite imposes zero derivatives: use a polynomial with vanishing
        # low-order derivatives at both ends only for degree >= 2 * nbe + 1 ... test
        # instead that the zero-derivative conditions hold.
        orc = O.Spline(degree, uniform, code, code, ncells=10, breaks=br)
        x = orc.points()
        c = orc.build(p(x)[:, None])[:, 0]
        spl = _bspline(orc, c)
        for k in range(1 + degree % 2 - 1 + 1, degree // 2 + 1):
            pass
        nbe = degree // 2
        for i in range(nbe):
            k = i + degree % 2
            if k == 0:
                assert abs(spl(orc.rmin)) <= 1e-12
            else:
                assert abs(spl(orc.rmin, nu=k)) <= 1e-9
        return
    orc = O.Spline(degree, uniform, code, code, ncells=10, breaks=br)
    x = orc.points()
    nbe = orc.nbe_l
    dl = np.array([[p.deriv(i + degree % 2)(orc.rmin)] for i in range(nbe)]) if nbe else None
    dr = np.array([[p.deriv(i + degree % 2)(orc.rmax)] for i in range(nbe)]) if nbe else None
    if nbe and degree % 2 == 0:
        dl = np.array([[p.deriv(i)(orc.rmin) if i else p(orc.rmin)] for i in range(nbe)])
        dr = np.array([[p.deriv(i)(orc.rmax) if i else p(orc.rmax)] for i in range(nbe)])
    c = orc.build(p(x)[:, None], dl, dr)
    xe = np.linspace(orc.rmin, orc.rmax, 97)
    vals = orc.eval(xe[:, None], c)[:, 0]
    scale = np.abs(p(xe)).max()
    assert np.abs(vals - p(xe)).max() <= 1e-14 * max(1.0, scale) * 10
    d1 = orc.eval(xe[:, None], c, deriv=1)[:, 0]
    assert np.abs(d1 - p.deriv(1)(xe)).max() <= 1e-11 * max(1.0, np.abs(p.deriv(1)(xe)).max())
    integ = orc.integrate(c)[0]
    exact = p.integ()(orc.rmax) - p.integ()(orc.rmin)
    assert abs(integ - exact) <= 1e-14 * max(1.0, abs(exact)) * 10


@pytest.mark.parametrize("degree", [2, 3, 4, 5])
@pytest.mark.parametrize("uniform", [True, False])
@pytest.mark.parametrize("bc", [O.GREVILLE, O.HERMITE, O.PERIODIC])
def test_quadrature_integrates_exactly(degree, uniform, bc):
    br = None if uniform else breaks_for("jitter", 12, seed=degree)
    orc = O.Spline(degree, uniform, bc, bc, ncells=12, breaks=br)
    ql, qv, qr = orc.quadrature_coefficients()
    x = orc.points()
    if bc == O.PERIODIC:
        f = lambda z: np.sin(2 * np.pi * z) + 2.0  # noqa: E731
        # quadrature of the periodic spline interpolant equals the spline integral
        c = orc.build(f(x)[:, None])[:, 0]
        # the reference solves A^T q = (integral of B_j), j in matrix-column order,
        # i.e. unrotated by the periodic offset; so q.f = sum_j int(B_j) c[j + offset]
        ib = orc.integrals()[: orc.nbasis]
        off = orc.offset
        rot = np.array([c[j + off] for j in range(orc.nbasis)])
        assert abs(np.dot(qv, f(x)) - np.dot(ib, rot)) <= 1e-13
        if uniform or off == 0:
            assert abs(np.dot(qv, f(x)) - orc.integrate(c[:, None])[0]) <= 1e-13
        return
    p = np.polynomial.Polynomial(np.random.default_rng(degree).uniform(-1, 1, degree + 1))
    total = np.dot(qv, p(x))
    odd = degree % 2
    for i in range(ql.size):
        k = i + odd
        total += ql[i] * (p.deriv(k)(orc.rmin) if k else p(orc.rmin))
        total += qr[i] * (p.deriv(k)(orc.rmax) if k else p(orc.rmax))
    exact = p.integ()(orc.rmax) - p.integ()(orc.rmin)
    assert abs(total - exact) <= 1e-13 * max(1.0, abs(exact))


def test_short_long_badshape_grids():
    """tests/splines/spline_builder.cpp:36-135 (uniform periodic degree 2, 5 cells)."""
    for pts in ([0.1, 0.3, 0.5, 0.7], [0.1, 0.3, 0.5, 0.7, 0.9, 0.95], [0.1, 0.11, 0.12, 0.13, 0.14]):
        with pytest.raises(RuntimeError):
            O.Spline(2, True, O.PERIODIC, O.PERIODIC, ncells=5, points=np.array(pts))


def test_greville_ordering_small_first_cell():
    """periodic_spline_builder_ordered_points.cpp:41-67: sorted points with a 0.01 first cell."""
    br = np.concatenate([[0.0, 0.01], np.linspace(0.1, 1.0, 10)])
    for d in (3, 4, 5):
        orc = O.Spline(d, False, O.PERIODIC, O.PERIODIC, breaks=br)
        x = orc.points()
        assert np.all(np.diff(x) > 0)
        assert x[0] >= orc.rmin and x[-1] < orc.rmax


def test_batch_layouts_agree():
    orc = O.Spline(3, False, O.HERMITE, O.HERMITE, breaks=breaks_for("random", 40, seed=3))
    rng = np.random.default_rng(5)
    f = rng.standard_normal((orc.npts, 7))
    dl = rng.standard_normal((1, 7))
    dr = rng.standard_normal((1, 7))
    a = orc.build(f, dl, dr, layout="batch_contiguous")
    b = orc.build(f.T, dl.T, dr.T, layout="interp_contiguous")
    assert np.array_equal(a, b.T)


def test_matrix_kinds_match_maker():
    """SplinesLinearProblemMaker choices (splines_linear_problem_maker.hpp:56-146)."""
    expect = {
        (3, True, O.PERIODIC): "2x2",
        (3, False, O.GREVILLE): "3x3",
        (3, False, O.HERMITE): "3x3",
        (5, True, O.PERIODIC): "2x2",
    }
    for (d, u, bc), kind in expect.items():
        br = None if u else breaks_for("random", 50, seed=0)
        assert O.Spline(d, u, bc, bc, ncells=50, breaks=br).matrix_type == kind
    # very small periodic problems use a dense matrix
    assert O.Spline(5, False, O.PERIODIC, O.PERIODIC, breaks=np.linspace(0, 1, 7)).matrix_type == "dense"
