"""Pins the oracle's basis, knots and cell search (CPU only).

Checks are independent of the oracle's formulas: scipy.interpolate.BSpline
(design matrix and derivatives), closed forms, partition of unity
(tests/splines/bsplines.cpp:67-137), edge rounding (bsplines.cpp:156-217) and
brute-force cell search.
"""
import numpy as np
import pytest
import scipy.interpolate as si

import oracle as O
from cases import breaks_for


def scipy_knots(orc):
    return orc.knots()


@pytest.mark.parametrize("degree", range(1, 7))
@pytest.mark.parametrize("ncells", [10, 20, 100])
@pytest.mark.parametrize("uniform", [True, False])
@pytest.mark.parametrize("periodic", [True, False])
def test_partition_of_unity(degree, ncells, uniform, periodic):
    bc = O.PERIODIC if periodic else O.GREVILLE
    br = None if uniform else breaks_for("jitter", ncells, seed=degree)
    orc = O.Spline(degree, uniform, bc, bc, ncells=ncells, breaks=br)
    xs = np.linspace(orc.rmin, orc.rmax, 1001)
    for x in xs:
        _, v = orc.eval_basis(x)
        assert abs(v.sum() - 1.0) <= 1e-15
        assert np.all(v >= -1e-15)


def test_cubic_values_at_knots():
    """Uniform cubic B-spline at a knot: (1/6, 2/3, 1/6, 0)."""
    orc = O.Spline(3, True, O.PERIODIC, O.PERIODIC, ncells=8)
    j, v = orc.eval_basis(orc.rmin + 3 * (orc.rmax - orc.rmin) / 8)
    np.testing.assert_allclose(v[:3], [1 / 6, 2 / 3, 1 / 6], rtol=0, atol=2e-16)
    assert abs(v[3]) <= 2e-16


@pytest.mark.parametrize("degree", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("uniform", [True, False])
def test_basis_vs_scipy(degree, uniform):
    ncells = 13
    br = None if uniform else breaks_for("random", ncells, seed=7)
    orc = O.Spline(degree, uniform, O.GREVILLE, O.GREVILLE, ncells=ncells, breaks=br)
    t = orc.knots()
    rng = np.random.default_rng(degree)
    xs = rng.uniform(orc.rmin, orc.rmax, 200)
    dm = si.BSpline.design_matrix(xs, t, degree).toarray()
    for n, x in enumerate(xs):
        j, v = orc.eval_basis(x)
        ref = dm[n, j: j + degree + 1]
        np.testing.assert_allclose(v, ref, rtol=0, atol=1e-14)
        # derivatives of order 1..degree against scipy
        jd, d = orc.eval_basis_and_n_derivs(x, degree)
        assert jd == j
        for k in range(1, degree + 1):
            for s in range(degree + 1):
                c = np.zeros(orc.nbasis)
                c[j + s] = 1.0
                refd = si.BSpline(t, c, degree)(x, nu=k)
                assert abs(d[s, k] - refd) <= 1e-9 * max(1.0, abs(refd)), (k, s)


@pytest.mark.parametrize("degree", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("uniform", [True, False])
def test_edge_rounding(degree, uniform):
    """bsplines.cpp:156-217: rmin - eps -> first B-spline, rmax + eps -> last."""
    ncells = 10
    br = None if uniform else np.array([0.0 + i * (0.2 / ncells) for i in range(ncells + 1)])
    orc = O.Spline(degree, uniform, O.GREVILLE, O.GREVILLE, ncells=ncells, rmin=0.0, rmax=0.2, breaks=br)
    eps = np.finfo(float).eps
    j0, _ = orc.eval_basis(0.0 - eps)
    j1, _ = orc.eval_basis(0.2 + eps)
    assert j0 == 0
    assert j1 == orc.size - 1 - degree


@pytest.mark.parametrize("degree", [2, 3, 5])
def test_nonuniform_cell_search_bruteforce(degree):
    br = breaks_for("random", 57, seed=11)
    orc = O.Spline(degree, False, O.GREVILLE, O.GREVILLE, breaks=br)
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(0, 1, 500), br, np.nextafter(br, 2), np.nextafter(br, -1)])
    xs = xs[(xs >= 0) & (xs <= 1)]
    for x in xs:
        j, _ = orc.eval_basis(x)
        # brute force: the cell is the last break point <= x, clamped to [0, ncells-1]
        cell = min(max(np.searchsorted(br, x, side="right") - 1, 0), br.size - 2)
        assert j == cell


@pytest.mark.parametrize("degree", [1, 3, 4])
def test_uniform_cell_bruteforce(degree):
    ncells = 37
    orc = O.Spline(degree, True, O.GREVILLE, O.GREVILLE, ncells=ncells, rmin=-1.3, rmax=2.9)
    t = orc.knots()
    br = t[degree: degree + ncells + 1]
    rng = np.random.default_rng(1)
    xs = rng.uniform(-1.3, 2.9, 2000)
    for x in xs:
        j, _ = orc.eval_basis(x)
        # the cell is the one whose [start, end] contains x (ties at an edge may go either way)
        assert br[j] - 1e-12 <= x <= br[j + 1] + 1e-12


def test_uniform_knots_closed_form():
    """init_ghosted: knot k = (rmin - d h) + k h with h = (rmax - rmin) / ncells."""
    d, n, a, b = 3, 20, 0.25, 1.75
    orc = O.Spline(d, True, O.PERIODIC, O.PERIODIC, ncells=n, rmin=a, rmax=b)
    h = (b - a) / n
    k = np.arange(n + 1 + 2 * d)
    assert np.array_equal(orc.knots(), (a - d * h) + k * h)


@pytest.mark.parametrize("periodic", [True, False])
def test_nonuniform_knots(periodic):
    d = 3
    br = breaks_for("random", 12, seed=2)
    bc = O.PERIODIC if periodic else O.GREVILLE
    orc = O.Spline(d, False, bc, bc, breaks=br)
    t = orc.knots()
    assert np.array_equal(t[d: d + br.size], br)
    L = br[-1] - br[0]
    for i in range(1, d + 1):
        if periodic:
            assert t[d - i] == br[-1 - i] - L
            assert t[d + 12 + i] == br[i] + L
        else:
            assert t[d - i] == br[0] and t[d + 12 + i] == br[-1]


@pytest.mark.parametrize("uniform", [True, False])
@pytest.mark.parametrize("degree", [1, 2, 3, 4, 5])
def test_integrals_closed_form(uniform, degree):
    ncells = 15
    br = None if uniform else breaks_for("random", ncells, seed=degree)
    for bc in (O.PERIODIC, O.GREVILLE):
        orc = O.Spline(degree, uniform, bc, bc, ncells=ncells, rmin=0.0, rmax=1.0, breaks=br)
        ib = orc.integrals()
        t = orc.knots()
        # the integral of B_i over [rmin, rmax] equals (t_{i+d+1} - t_i)/(d+1) when the
        # support lies inside the domain; integrate numerically with scipy otherwise
        for i in range(orc.nbasis):
            c = np.zeros(orc.size)
            c[i] = 1.0
            if bc == O.PERIODIC and i < degree:
                c[orc.nbasis + i] = 1.0  # periodic repeats
            spl = si.BSpline(t, c, degree, extrapolate=False)
            ref = spl.integrate(orc.rmin, orc.rmax)
            if bc == O.PERIODIC:
                # a periodic B-spline integrates to the length of its support / (d+1)
                ref = (t[i + degree + 1] - t[i]) / (degree + 1)
            assert abs(ib[i] - ref) <= 1e-14, (i, ib[i], ref)
        if bc == O.PERIODIC:
            assert np.all(ib[orc.nbasis:] == 0.0)
        assert abs(ib[: orc.nbasis].sum() - 1.0) <= 1e-13
