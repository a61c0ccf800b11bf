"""Pins the oracle's N-dimensional restatement (splinend_build, splinend_eval,
splinend_integrate; SplineBuilder3D / SplineEvaluatorND) without the
reference, by what the mathematics fixes:
- one dimension is the 1D oracle, bit for bit;
- two dimensions agree with the 2D oracle (which associates the products the
  other way) to rounding;
- a separable coefficient block c = a (x) b (x) c gives the product of the 1D
  evaluations, derivatives and integrals (catches transposed strides and
  swapped dimensions, since every dimension has its own size and degree);
- tensor polynomials of degree <= d in each dimension are reproduced exactly
  by the 3D interpolation, and their integrals are the closed forms;
- wrap and Null rules per dimension.
"""
import numpy as np
import pytest

import oracle as O
from cases import G, P, breaks_for


def _dims():
    """Three different 1D bases: non-uniform cubic (Greville), uniform
    periodic quadratic, uniform quartic (Greville)."""
    s1 = O.Spline(3, False, G, G, breaks=breaks_for("random", 9, seed=1))
    s2 = O.Spline(2, True, P, P, ncells=7, rmin=0.0, rmax=1.0)
    s3 = O.Spline(4, True, G, G, ncells=6, rmin=-1.0, rmax=2.0)
    return [s1, s2, s3]


def _eval1(s, x, c, deriv=0):
    """1D oracle evaluation of one coefficient vector at points x."""
    return s.eval(x[:, None], c[:, None], deriv=deriv)[:, 0]


def test_nd_one_dimension_is_the_1d_oracle():
    rng = np.random.default_rng(0)
    for s in _dims():
        coef = rng.standard_normal((3, s.size))
        x = rng.uniform(s.rmin - 0.2, s.rmax + 0.2, (3, 40))
        for d in range(0, min(s.degree, 3) + 1):
            xin = x if d == 0 else np.clip(x, s.rmin, s.rmax)
            got = O.splinend_eval([s], xin[..., None], coef, orders=[d])
            ref = np.stack([_eval1(s, xin[l], coef[l], d) for l in range(3)])
            assert np.array_equal(got, ref), d


def test_nd_two_dimensions_matches_2d_oracle():
    s1, s2, _ = _dims()
    rng = np.random.default_rng(1)
    coef = rng.standard_normal((2, s1.size, s2.size))
    coords = np.stack([rng.uniform(0, 1, (2, 60)), rng.uniform(-0.5, 1.5, (2, 60))], axis=-1)
    for orders in [(0, 0), (1, 0), (0, 2), (1, 1)]:
        ref = O.spline2d_eval(s1, s2, coords, coef, *orders)
        got = O.splinend_eval([s1, s2], coords, coef, orders=list(orders))
        assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max(), orders


@pytest.mark.parametrize("orders", [(0, 0, 0), (1, 0, 0), (0, 2, 0), (0, 0, 3), (1, 1, 1), (2, 1, 3)])
def test_nd_separable_product(orders):
    ss = _dims()
    rng = np.random.default_rng(2)
    vecs = [rng.standard_normal(s.size) for s in ss]
    # the periodic dimension repeats its first `degree` coefficients
    vecs[1][ss[1].nbasis:] = vecs[1][: ss[1].degree]
    coef = np.einsum("i,j,k->ijk", *vecs)[None]
    pts = [rng.uniform(s.rmin, s.rmax, 50) for s in ss]
    coords = np.stack(pts, axis=-1)[None]
    got = O.splinend_eval(ss, coords, coef, orders=list(orders))[0]
    ref = np.prod([_eval1(s, p, v, d) for s, p, v, d in zip(ss, pts, vecs, orders)], axis=0)
    assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())
    # integrals: product of the 1D integrals
    integ = O.splinend_integrate(ss, coef)[0]
    ref_i = np.prod([s.integrate(v[:, None])[0] for s, v in zip(ss, vecs)])
    assert abs(integ - ref_i) <= 1e-13 * max(1.0, abs(ref_i))


def test_nd_rules_wrap_and_null():
    ss = _dims()
    rng = np.random.default_rng(3)
    vecs = [rng.standard_normal(s.size) for s in ss]
    vecs[1][ss[1].nbasis:] = vecs[1][: ss[1].degree]
    coef = np.einsum("i,j,k->ijk", *vecs)[None]
    inside = np.array([[0.3, 0.4, 0.5]])
    outside = inside.copy()
    outside[0, 1] += 3.0  # periodic dimension: wraps to the same point
    a = O.splinend_eval(ss, inside[None], coef)[0, 0]
    b = O.splinend_eval(ss, outside[None], coef)[0, 0]
    assert abs(a - b) <= 1e-13 * abs(a)
    for k, delta in [(0, -0.1), (0, 1.1), (2, -1.5), (2, 2.5)]:
        p = inside.copy()
        p[0, k] = delta
        assert O.splinend_eval(ss, p[None], coef)[0, 0] == 0.0


def test_3d_build_reproduces_tensor_polynomials():
    """Interpolation of P(x) Q(y) R(z), each of degree <= the dimension's
    spline degree, is exact (non-periodic Greville closures), so evaluation
    anywhere returns the polynomial and the integral its closed form."""
    s1 = O.Spline(3, False, G, G, breaks=breaks_for("random", 8, seed=4))
    s2 = O.Spline(2, True, G, G, ncells=5, rmin=0.0, rmax=1.0)
    s3 = O.Spline(4, False, G, G, breaks=breaks_for("jitter", 6, seed=5))
    ss = [s1, s2, s3]
    polys = [np.polynomial.Polynomial([0.3, -1.2, 0.7, 0.5]), np.polynomial.Polynomial([1.1, 0.4, -0.9]),
             np.polynomial.Polynomial([0.2, 0.1, -0.3, 0.8, -0.6])]
    g = [s.points() for s in ss]
    vals = np.einsum("i,j,k->ijk", *[p(x) for p, x in zip(polys, g)])[None]
    coef = O.splinend_build(ss, vals)
    assert coef.shape == (1, s1.size, s2.size, s3.size)
    rng = np.random.default_rng(6)
    coords = np.stack([rng.uniform(s.rmin, s.rmax, 200) for s in ss], axis=-1)[None]
    ref = np.prod([p(coords[0, :, k]) for k, p in enumerate(polys)], axis=0)
    got = O.splinend_eval(ss, coords, coef)[0]
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    # derivative in every dimension
    dref = np.prod([p.deriv()(coords[0, :, k]) for k, p in enumerate(polys)], axis=0)
    dgot = O.splinend_eval(ss, coords, coef, orders=[1, 1, 1])[0]
    assert np.abs(dgot - dref).max() <= 1e-10 * np.abs(dref).max()
    integ = O.splinend_integrate(ss, coef)[0]
    ref_i = np.prod([p.integ()(s.rmax) - p.integ()(s.rmin) for p, s in zip(polys, ss)])
    assert abs(integ - ref_i) <= 1e-12 * abs(ref_i)


def test_3d_build_is_three_1d_builds():
    """Each 1D pass of splinend_build is the 1D oracle build along that axis."""
    ss = _dims()
    rng = np.random.default_rng(7)
    vals = rng.standard_normal((2,) + tuple(s.npts for s in ss))
    got = O.splinend_build(ss, vals)
    cur = vals
    for ax, s in enumerate(ss, start=1):
        moved = np.moveaxis(cur, ax, 0)
        flat = moved.reshape(moved.shape[0], -1)
        out = s.build(flat).reshape((s.size,) + moved.shape[1:])
        cur = np.moveaxis(out, 0, ax)
    assert np.array_equal(got, cur)


def hermite_3d_data(ss, polys, nb=1):
    """Interpolation data of P1(x) P2(y) P3(z) for the 3D builder with the
    reference's argument names: values, boundary derivatives per dimension
    (row k is order k + degree % 2), and the cross derivatives on edges and
    corners, each taken from the polynomials' exact derivatives."""
    pts = [s.points() for s in ss]
    bnd = [{"min": s.rmin, "max": s.rmax} for s in ss]

    def dv(k, p, order, t):
        q = polys[k].deriv(order) if order else polys[k]
        return q(t)

    def field(spec):
        # spec per dimension: None (the interpolation points) or (side, orders list)
        axes = []
        for k in range(3):
            if spec[k] is None:
                axes.append(np.array([dv(k, None, 0, t) for t in pts[k]]))
            else:
                side, ords = spec[k]
                axes.append(np.array([dv(k, None, o, bnd[k][side]) for o in ords]))
        a = np.einsum("i,j,k->ijk", *axes)
        return np.repeat(a[None], nb, axis=0)

    ords = [[i + s.degree % 2 for i in range(s.nbv_l)] for s in ss]
    out = {"vals": field([None, None, None])}
    for k in range(3):
        for side in ("min", "max"):
            spec = [None, None, None]
            spec[k] = (side, ords[k])
            out[f"derivs_{side}{k + 1}"] = field(spec)
    for (a, b) in ((0, 1), (1, 2), (0, 2)):
        for sa in ("min", "max"):
            for sb in ("min", "max"):
                spec = [None, None, None]
                spec[a], spec[b] = (sa, ords[a]), (sb, ords[b])
                out[f"mixed_derivs_{sa}{a + 1}_{sb}{b + 1}"] = field(spec)
    for s1 in ("min", "max"):
        for s2 in ("min", "max"):
            for s3 in ("min", "max"):
                out[f"mixed_derivs_{s1}1_{s2}2_{s3}3"] = field([(s1, ords[0]), (s2, ords[1]), (s3, ords[2])])
    vals = out.pop("vals")
    return vals, out


def test_3d_hermite_build_reproduces_tensor_polynomials():
    """SplineBuilder3D with Hermite closures in all three dimensions (boundary
    derivatives, edge and corner cross derivatives, spline_builder_3d.hpp:
    889-1090) reproduces P1(x) P2(y) P3(z) of degree <= each spline degree
    exactly; the triple corner derivatives are used (changing one changes
    the coefficients), and an all-Greville build ignores the derivative data."""
    from cases import breaks_for

    H, G = O.HERMITE, O.GREVILLE
    ss = [O.Spline(3, False, H, H, breaks=breaks_for("random", 6, seed=1)),
          O.Spline(5, True, H, H, ncells=5, rmin=-0.5, rmax=1.5),
          O.Spline(3, True, H, H, ncells=7, rmin=0.0, rmax=2.0)]
    polys = [np.polynomial.Polynomial([0.4, -1.1, 0.6, 0.9]),
             np.polynomial.Polynomial([1.0, 0.3, -0.7, 0.2, 0.5, -0.4]),
             np.polynomial.Polynomial([-0.2, 0.8, 0.5, -0.3])]
    vals, d = hermite_3d_data(ss, polys)
    coef = O.spline3d_build(*ss, vals, d)
    rng = np.random.default_rng(4)
    xyz = np.stack([rng.uniform(s.rmin, s.rmax, 400) for s in ss], axis=-1)[None]
    ref = polys[0](xyz[0, :, 0]) * polys[1](xyz[0, :, 1]) * polys[2](xyz[0, :, 2])
    got = O.splinend_eval(ss, xyz, coef)[0]
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    dref = polys[0].deriv()(xyz[0, :, 0]) * polys[1].deriv(2)(xyz[0, :, 1]) * polys[2](xyz[0, :, 2])
    dgot = O.splinend_eval(ss, xyz, coef, orders=[1, 2, 0])[0]
    assert np.abs(dgot - dref).max() <= 1e-10 * np.abs(dref).max()
    bad = dict(d)
    bad["mixed_derivs_max1_min2_max3"] = d["mixed_derivs_max1_min2_max3"].copy()
    bad["mixed_derivs_max1_min2_max3"][0, 0, 0, 0] += 1.0
    assert np.abs(O.spline3d_build(*ss, vals, bad) - coef).max() > 1e-6
    gs = [O.Spline(3, False, G, G, breaks=breaks_for("random", 6, seed=1)),
          O.Spline(3, True, G, G, ncells=5, rmin=-0.5, rmax=1.5),
          O.Spline(3, True, G, G, ncells=7, rmin=0.0, rmax=2.0)]
    gv = np.random.default_rng(5).standard_normal((2,) + tuple(s.npts for s in gs))
    assert np.array_equal(O.spline3d_build(*gs, gv, {}), O.splinend_build(gs, gv))
