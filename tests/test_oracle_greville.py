"""Pins the oracle's Greville interpolation points (CPU only).

GrevilleInterpolationPoints (greville_interpolation_points.hpp:46-295) is
checked here against properties that do not go through the oracle's
knot-averaging code:

- Linear precision.  The Greville abscissae xi_j are exactly the B-spline
  coefficients of the identity: sum_j xi_j B_j(x) = x.  For non-periodic
  bases with Greville closures the point set is the set of Greville
  abscissae of the clamped knot vector (break points, end knots repeated
  degree + 1 times), so scipy's BSpline with those knots and the points as
  coefficients must reproduce x.  For periodic bases the sorted points are
  the abscissae of the periodic extension after a rotation; some rotation
  (at most `degree` points moved by one period) must make the extended
  coefficients reproduce x.
- Mirror symmetry.  On a mesh symmetric about the middle of the domain, the
  points of (closure A, closure B) are the reflection of those of (B, A).
  This fixes which Greville abscissae the Hermite closures drop.
- Closed forms for uniform bases: shifted uniform points
  rmin + (shift + i) dx (shift 1/2 for even degree) and the Greville-closure
  edge points rmin + dx i (i + 1) / (2 degree), the clamped-knot abscissae
  (degree - i) rmin + (rmin + dx) + ... + (rmin + i dx), divided by degree.

Each check is also run on a deliberately wrong point set (averages of
degree + 1 knots; a start knot off by one) to show that it rejects it.
"""
import numpy as np
import pytest
import scipy.interpolate as si

import oracle as O
from cases import breaks_for

G, H, P = O.GREVILLE, O.HERMITE, O.PERIODIC


def clamped_knots(breaks, degree):
    return np.concatenate([[breaks[0]] * degree, breaks, [breaks[-1]] * degree])


def linear_precision_error(points, breaks, degree):
    """max |sum_j p_j B_j(x) - x| over [rmin, rmax] with clamped knots."""
    t = clamped_knots(breaks, degree)
    if points.size != t.size - degree - 1:
        return np.inf
    x = np.concatenate([np.linspace(breaks[0], breaks[-1], 2001), breaks])
    spl = si.BSpline(t, points, degree, extrapolate=False)
    return float(np.max(np.abs(spl(x) - x)))


def periodic_linear_precision_error(points, knots, degree):
    """Smallest max |f(x) - x| over the rotations r = 0..degree, where f has
    the periodic extension's coefficients: the first r abscissae are the last
    r points minus one period, the rest follow in order, and coefficients
    past nbasis repeat with one period added (knots: the ghosted knot
    vector, ncells + 1 + 2 degree entries)."""
    nb = points.size
    rmin, rmax = knots[degree], knots[-degree - 1]
    length = rmax - rmin
    # the points are increasing and inside [rmin, rmax]; rmax itself occurs
    # when an abscissa that is 0 in exact arithmetic rounds below rmin and is
    # moved to the end with one period added
    if not (np.all(np.diff(points) > 0) and points[0] >= rmin and points[-1] <= rmax):
        return np.inf
    x = np.linspace(rmin, rmax, 2001)[:-1]
    best = np.inf
    for r in range(degree + 1):
        xi = np.concatenate([points[nb - r:] - length, points[: nb - r]])
        c = np.concatenate([xi, xi[:degree] + length])
        spl = si.BSpline(knots, c, degree, extrapolate=False)
        best = min(best, float(np.max(np.abs(spl(x) - x))))
    return best


def oracle_points(degree, bcl, bcu, breaks=None, ncells=None):
    if breaks is None:
        return O.Spline(degree, True, bcl, bcu, ncells=ncells, rmin=0.0, rmax=1.0)
    return O.Spline(degree, False, bcl, bcu, breaks=breaks)


MESHES = ["uniform", "random", "geometric", "jitter"]


@pytest.mark.parametrize("degree", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("mesh", MESHES)
def test_nonuniform_greville_linear_precision(degree, mesh):
    br = breaks_for(mesh, 23, seed=degree + 7, ratio=1.2)
    orc = oracle_points(degree, G, G, breaks=br)
    p = orc.points()
    assert linear_precision_error(p, br, degree) <= 1e-14
    # a mean of degree + 1 knots, or one starting a knot late, is rejected
    t = clamped_knots(br, degree)
    nb = p.size
    wrong1 = np.array([t[j + 1: j + degree + 2].mean() for j in range(nb)])
    wrong2 = np.array([t[j + 2: j + degree + 2].mean() for j in range(nb)])
    assert linear_precision_error(wrong1, br, degree) > 1e-6
    assert linear_precision_error(wrong2, br, degree) > 1e-6


@pytest.mark.parametrize("degree", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("ncells", [5, 12, 31])
def test_uniform_greville_closure(degree, ncells):
    """Uniform bases, Greville at both ends: the closed form of the edge
    points and, up to degree 4, the clamped-knot abscissae.  For degree 5 the
    reference keeps the uniform points next to the edge points (abscissa 3
    is rmin + dx, where the clamped knots give rmin + 6 dx / 5), so linear
    precision does not hold there and the closed forms pin the points."""
    orc = oracle_points(degree, G, G, ncells=ncells)
    p = orc.points()
    br = np.linspace(0.0, 1.0, ncells + 1)
    if degree <= 4:
        assert linear_precision_error(p, br, degree) <= 1e-14
    else:
        assert linear_precision_error(p, br, degree) > 1e-3
    dx = 1.0 / ncells
    for i in range(degree // 2 + 1):
        assert abs(p[i] - dx * i * (i + 1) / (2 * degree)) <= 4e-16
        assert abs(p[-1 - i] - (1.0 - dx * i * (i + 1) / (2 * degree))) <= 4e-16
    # the (degree + 1)-knot average of the edge points is rejected
    wrong = p.copy()
    for i in range(degree // 2 + 1):
        wrong[i] = dx * i * (i + 1) / (2 * (degree + 1))
    if 1 < degree <= 4:
        assert linear_precision_error(wrong, br, degree) > 1e-6
    np.testing.assert_allclose(p, expected_uniform(degree, ncells, G, G), rtol=0, atol=4.5e-16)


def expected_uniform(degree, ncells, bcl, bcu):
    """Closed form of the uniform points for every closure pair
    (greville_interpolation_points.hpp:46-69, :171-248), on [0, 1]."""
    dx = 1.0 / ncells
    shift = 0.5 if degree % 2 == 0 else 0.0
    nbe = lambda bc: degree // 2 if bc in (H, O.HOMOGENEOUS_HERMITE) else 0  # noqa: E731
    nl, nu = nbe(bcl), nbe(bcu)
    npts = ncells + degree - nl - nu
    if nl and nu:
        return (shift + np.arange(npts)) * dx
    p = np.full(npts, np.nan)
    if bcl == G:
        for i in range(degree // 2 + 1):
            p[i] = dx * i * (i + 1) / (2 * degree)
        start = degree // 2 + 1
    else:
        p[0] = shift * dx
        start = 1
    for k in range(ncells - 1):
        p[start + k] = (shift + 1 + k) * dx
    if bcu == G:
        for i in range(degree // 2 + 1):
            # for even degrees the last edge point lands on the slot of the
            # last interior point and replaces it, as in the reference
            p[npts - 1 - i] = 1.0 - dx * i * (i + 1) / (2 * degree)
    else:
        p[npts - 1] = (shift + ncells - 1 + degree % 2) * dx
    assert not np.any(np.isnan(p))
    return p


CLOSURE_PAIRS = [(G, G), (H, H), (H, G), (G, H), (O.HOMOGENEOUS_HERMITE, G), (G, O.HOMOGENEOUS_HERMITE),
                 (O.HOMOGENEOUS_HERMITE, O.HOMOGENEOUS_HERMITE)]


@pytest.mark.parametrize("degree", [2, 3, 4, 5])
@pytest.mark.parametrize("bcs", CLOSURE_PAIRS)
@pytest.mark.parametrize("ncells", [6, 17])
def test_uniform_closure_pairs(degree, bcs, ncells):
    p = oracle_points(degree, bcs[0], bcs[1], ncells=ncells).points()
    np.testing.assert_allclose(p, expected_uniform(degree, ncells, *bcs), rtol=0, atol=4.5e-16)


def symmetric_breaks(ncells, seed):
    """Random break points on [0, 1] with x_k = 1 - x_(ncells - k) exactly
    (multiples of 2^-20, so 1 - x rounds to nothing)."""
    rng = np.random.default_rng(seed)
    m = ncells // 2
    u = np.round(rng.uniform(0.0, 0.5, m - 1 + ncells % 2) * 2.0 ** 20) / 2.0 ** 20
    left = np.concatenate([[0.0], np.sort(u)])
    mid = [0.5] if ncells % 2 == 0 else []
    return np.concatenate([left, mid, (1.0 - left)[::-1]])


@pytest.mark.parametrize("degree", [2, 3, 4, 5])
@pytest.mark.parametrize("uniform", [True, False])
@pytest.mark.parametrize("bcs", [(H, G), (H, H), (O.HOMOGENEOUS_HERMITE, G)])
@pytest.mark.parametrize("ncells", [13, 14])
def test_mirror_symmetry(degree, uniform, bcs, ncells):
    """(A, B) on a symmetric mesh reflects to (B, A): fixes which end the
    Hermite closures drop their abscissae from."""
    br = None
    if not uniform:
        br = symmetric_breaks(ncells, seed=degree)
        assert br.size == ncells + 1 and np.array_equal(br, 1.0 - br[::-1])
    a = oracle_points(degree, bcs[0], bcs[1], breaks=br, ncells=ncells).points()
    b = oracle_points(degree, bcs[1], bcs[0], breaks=br, ncells=ncells).points()
    np.testing.assert_allclose(a, 1.0 - b[::-1], rtol=0, atol=1e-15)


@pytest.mark.parametrize("degree", [2, 3, 4, 5])
@pytest.mark.parametrize("bcs", [(H, G), (H, H), (G, H)])
def test_nonuniform_hermite_drops_end_abscissae(degree, bcs):
    """With Hermite closures the points are the Greville-closure points (pinned
    by linear precision above) without the first / last degree // 2."""
    br = breaks_for("random", 19, seed=degree)
    full = oracle_points(degree, G, G, breaks=br).points()
    p = oracle_points(degree, bcs[0], bcs[1], breaks=br).points()
    nl = degree // 2 if bcs[0] == H else 0
    nu = degree // 2 if bcs[1] == H else 0
    np.testing.assert_array_equal(p, full[nl: full.size - nu])


@pytest.mark.parametrize("degree", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("mesh", ["uniform-basis", "uniform", "random", "geometric"])
def test_periodic_linear_precision(degree, mesh):
    ncells = 21
    if mesh == "uniform-basis":
        orc = oracle_points(degree, P, P, ncells=ncells)
    else:
        orc = oracle_points(degree, P, P, breaks=breaks_for(mesh, ncells, seed=degree, ratio=1.15))
    p = orc.points()
    t = orc.knots()
    assert p.size == ncells
    assert periodic_linear_precision_error(p, t, degree) <= 1e-14
    if orc.uniform:
        shift = 0.5 if degree % 2 == 0 else 0.0
        np.testing.assert_allclose(p, (shift + np.arange(ncells)) / ncells, rtol=0, atol=4.5e-16)
    # averages of degree + 1 knots (sorted into the domain) are rejected
    length = orc.rmax - orc.rmin
    wrong = np.sort((np.array([t[j + 1: j + degree + 2].mean() for j in range(ncells)]) - orc.rmin) % length
                    + orc.rmin)
    assert periodic_linear_precision_error(wrong, t, degree) > 1e-6
