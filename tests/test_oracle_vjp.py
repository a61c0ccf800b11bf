"""Pins the oracle's adjoints (CPU only).

- build_vjp against the Jacobian assembled from oracle builds applied to unit
  vectors (values and Hermite derivatives), and the identity
  <build(u), g> = <u, build_vjp(g)>.
- eval_vjp: coefficient gradient against the Jacobian of eval in the
  coefficients (unit vectors), coordinate gradient against central finite
  differences (h = 1e-6) and against deriv(1).
"""
import numpy as np
import pytest

import oracle as O
from cases import Case, G, H, HH, P, breaks_for

CASES = [Case(3, True, P, ncells=12), Case(5, True, P, ncells=12), Case(3, False, G, ncells=12),
         Case(3, False, H, ncells=12), Case(5, False, H, ncells=12), Case(4, True, G, H, ncells=12),
         Case(5, False, HH, ncells=12), Case(2, False, P, ncells=12, mesh="jitter")]


@pytest.mark.parametrize("case", CASES, ids=lambda c: c.id)
def test_build_vjp_equals_jacobian_transpose(case):
    orc = case.oracle()
    m, nl, nr = orc.npts, orc.nbv_l, orc.nbv_r
    nin = m + nl + nr
    J = np.zeros((orc.size, nin))
    for k in range(nin):
        u = np.zeros(nin)
        u[k] = 1.0
        dl = u[m: m + nl][:, None] if nl else None
        dr = u[m + nl:][:, None] if nr else None
        if case.bcl == H and not nl:
            dl = np.zeros((0, 1))
        J[:, k] = orc.build(u[:m][:, None], dl, dr)[:, 0]
    rng = np.random.default_rng(3)
    g = rng.standard_normal((orc.size, 2))
    gv, gdl, gdr = orc.build_vjp(g)
    ref = J.T @ g
    got = np.concatenate([gv, gdl, gdr], axis=0)
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("case", CASES, ids=lambda c: c.id)
def test_eval_vjp_coefficients_and_fd(case):
    orc = case.oracle()
    rng = np.random.default_rng(4)
    coef = rng.standard_normal((orc.size, 1))
    if case.periodic:
        coef[orc.nbasis:] = coef[: case.degree]
    x = rng.uniform(orc.rmin, orc.rmax, (40, 1))
    w = rng.standard_normal((40, 1))
    gc, gx = orc.eval_vjp(x, w, coef)
    # Jacobian in the coefficients, column by column
    J = np.zeros((40, orc.size))
    for j in range(orc.size):
        e = np.zeros((orc.size, 1))
        e[j] = 1.0
        J[:, j] = orc.eval(x, e)[:, 0]
    assert np.abs(gc[:, 0] - J.T @ w[:, 0]).max() <= 1e-14 * max(1.0, np.abs(J.T @ w[:, 0]).max())
    # coordinates: central differences, h = 1e-6, and deriv(1)
    h = 1e-6
    fd = (orc.eval(x + h, coef) - orc.eval(x - h, coef)) / (2 * h)
    inside = (x > orc.rmin + 2 * h) & (x < orc.rmax - 2 * h)
    assert np.abs((gx - w * fd)[inside]).max() <= 1e-6 * max(1.0, np.abs(w * fd).max())
    assert np.abs(gx - w * orc.eval(x, coef, deriv=1)).max() <= 1e-13 * max(1.0, np.abs(gx).max())


def test_eval_vjp_outside_domain():
    orc = O.Spline(3, False, G, G, breaks=breaks_for("random", 10, seed=1))
    coef = np.random.default_rng(0).standard_normal((orc.size, 1))
    x = np.array([[-0.5], [1.5]])
    w = np.ones((2, 1))
    gc, gx = orc.eval_vjp(x, w, coef, extrap=("null", "constant"))
    assert np.all(gx == 0.0)
    # null rule: no coefficient gradient; constant rule: the basis at rmax
    _, v = orc.eval_basis(orc.rmax)
    assert np.array_equal(gc[-4:, 0], v) and np.all(gc[:-4] == 0.0)
