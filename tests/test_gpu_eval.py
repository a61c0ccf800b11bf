"""GPU parity of SplineEvaluator (operator(), deriv(), integrate()) against the oracle.

Bars: cell indices are bit-exact (checked through values at break points and
near cell edges); values <= 1e-13 relative error (max norm).
"""
import numpy as np
import pytest

from cases import ALL_CASES, Case, G, H, P, values
from conftest import rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-13


def _coef(case, nbatch, seed=0):
    orc = case.oracle()
    rng = np.random.default_rng(seed)
    coef = rng.standard_normal((orc.size, nbatch))
    if case.periodic:  # periodic coefficients repeat the first `degree` ones
        coef[orc.nbasis:] = coef[: case.degree]
    return orc, coef


def _points(orc, npts, nbatch, seed, outside=0.0):
    rng = np.random.default_rng(seed)
    lo, hi = orc.rmin, orc.rmax
    L = hi - lo
    x = rng.uniform(lo - outside * L, hi + outside * L, (npts, nbatch))
    # exact break points and their neighbours exercise the cell-edge rounding
    br = orc.knots()[orc.degree: orc.degree + orc.ncells + 1]
    k = min(npts, br.size)
    x[:k, 0] = br[:k]
    x[:k, -1] = np.nextafter(br[:k], np.inf)
    x[:k, nbatch // 2] = np.nextafter(br[:k], -np.inf)
    return x


@pytest.mark.parametrize("case", ALL_CASES, ids=lambda c: c.id)
@pytest.mark.parametrize("layout", ["batch_contiguous", "interp_contiguous"])
def test_eval_and_deriv(case, layout, cuda):
    import torch
    import ddc_amd as D

    orc, coef = _coef(case, 9)
    bs, _ = case.product()
    x = _points(orc, 50, 9, 1, outside=0.2)
    rules = ("constant", "null")
    ev = D.SplineEvaluator(bs, D.ConstantExtrapolationRule(), D.NullExtrapolationRule())
    for order in range(0, case.degree + 1):
        if order > 0:
            xin = np.clip(x, orc.rmin, orc.rmax)
        else:
            xin = x
        ref = orc.eval(xin, coef, deriv=order, extrap=rules)
        if layout == "batch_contiguous":
            tx = torch.from_numpy(xin).to(cuda)
            tc = torch.from_numpy(coef).to(cuda)
            out = torch.empty_like(tx)
            (ev(out, tx, tc) if order == 0 else ev.deriv(order, out, tx, tc))
            got = out.cpu().numpy()
        else:
            tx = torch.from_numpy(np.ascontiguousarray(xin.T)).to(cuda)
            tc = torch.from_numpy(np.ascontiguousarray(coef.T)).to(cuda)
            out = torch.empty_like(tx)
            (ev(out, tx, tc, axis=1) if order == 0 else ev.deriv(order, out, tx, tc, axis=1))
            got = out.cpu().numpy().T
        scale = max(1.0, np.abs(ref).max())
        assert np.abs(got - ref).max() / scale <= TOL * (10 ** order), f"order {order}"


@pytest.mark.parametrize("case", [c for c in ALL_CASES if c.degree in (3, 5)], ids=lambda c: c.id)
def test_integrate(case, cuda):
    import torch
    import ddc_amd as D

    orc, coef = _coef(case, 11, seed=2)
    bs, _ = case.product()
    ev = D.SplineEvaluator(bs)
    ref = orc.integrate(coef)
    out = torch.empty(11, dtype=torch.float64, device=cuda)
    ev.integrate(out, torch.from_numpy(coef).to(cuda))
    assert rel_err(out.cpu().numpy(), ref) <= TOL


def test_build_then_eval_roundtrip(cuda):
    """Evaluating the built spline at the interpolation points returns the values."""
    import torch
    import ddc_amd as D

    case = Case(3, False, G, ncells=200, mesh="random")
    bs, builder = case.product()
    x = builder.interpolation_points()
    f = values(x.size, 64, 5, x=x)
    tv = torch.from_numpy(f).to(cuda)
    coef = torch.empty((bs.size(), 64), dtype=torch.float64, device=cuda)
    builder(coef, tv)
    ev = D.SplineEvaluator(bs)
    tx = torch.from_numpy(np.repeat(x[:, None], 64, axis=1)).to(cuda)
    out = torch.empty_like(tx)
    ev(out, tx, coef)
    assert rel_err(out.cpu().numpy(), f) <= 1e-13


def test_periodicity(cuda):
    """periodicity_spline_builder.cpp:170-202: f(x + 1.5) on an even function."""
    import torch
    import ddc_amd as D

    case = Case(3, True, P, ncells=64)
    orc = case.oracle()
    bs, builder = case.product()
    x = builder.interpolation_points()
    f = np.cos(2 * np.pi * x)[:, None]
    coef = torch.empty((bs.size(), 1), dtype=torch.float64, device=cuda)
    builder(coef, torch.from_numpy(f).to(cuda))
    ev = D.SplineEvaluator(bs)
    xe = np.linspace(0, 1, 101)[:, None] + 1.5
    out = torch.empty((101, 1), dtype=torch.float64, device=cuda)
    ev(out, torch.from_numpy(xe).to(cuda), coef)
    ref = orc.eval(xe, coef.cpu().numpy())
    assert rel_err(out.cpu().numpy(), ref) <= TOL
    assert np.abs(out.cpu().numpy()[:, 0] + np.cos(2 * np.pi * (xe[:, 0] - 1.5))).max() < 1e-5


@pytest.mark.parametrize("case", [Case(3, True, P, ncells=1024), Case(5, True, P, ncells=300),
                                  Case(3, False, G, ncells=1024, mesh="random"),
                                  Case(5, False, H, ncells=2000, mesh="random"),
                                  Case(4, True, G, ncells=100)], ids=lambda c: c.id)
@pytest.mark.parametrize("layout", ["batch_contiguous", "interp_contiguous"])
def test_eval_tiled(case, layout, cuda):
    """LDS-staged evaluator (>= 64 points, >= 8 lines), both thread mappings."""
    import torch
    import ddc_amd as D

    nbatch = 37
    orc, coef = _coef(case, nbatch, seed=4)
    bs, _ = case.product()
    x = _points(orc, 300, nbatch, 5, outside=0.1)
    ev = D.SplineEvaluator(bs, D.ConstantExtrapolationRule(), D.NullExtrapolationRule())
    bc = layout == "batch_contiguous"
    ax = 0 if bc else 1
    tc = torch.from_numpy(coef if bc else np.ascontiguousarray(coef.T)).to(cuda)
    for order in range(0, min(case.degree, 3) + 1):
        xin = x if order == 0 else np.clip(x, orc.rmin, orc.rmax)
        ref = orc.eval(xin, coef, deriv=order, extrap=("constant", "null"))
        tx = torch.from_numpy(xin if bc else np.ascontiguousarray(xin.T)).to(cuda)
        out = torch.empty_like(tx)
        (ev(out, tx, tc, axis=ax) if order == 0 else ev.deriv(order, out, tx, tc, axis=ax))
        got = out.cpu().numpy() if bc else out.cpu().numpy().T
        scale = max(1.0, np.abs(ref).max())
        assert np.abs(got - ref).max() / scale <= TOL * (10 ** order), f"order {order}"


@pytest.mark.parametrize("case", [Case(3, True, P, ncells=1024), Case(5, True, P, ncells=300),
                                  Case(1, True, G, ncells=200), Case(4, True, G, ncells=100)],
                         ids=lambda c: c.id)
@pytest.mark.parametrize("npts", [300, 1000])
def test_eval_stream(case, npts, cuda):
    """Persistent evaluator for uniform bases on large batch-contiguous batches
    (at least two 16-line tiles per CU), values and deriv(1): 8200 lines, so
    the last tile is ragged, and point counts that leave the last stage of
    each tile partly empty; extrapolated points on both sides.  Checked on
    192 lines spread over the batch (both ends included)."""
    import torch
    import ddc_amd as D

    nbatch = 8200
    orc, coef = _coef(case, nbatch, seed=6)
    bs, _ = case.product()
    x = _points(orc, npts, nbatch, 7, outside=0.1)
    ev = D.SplineEvaluator(bs, D.ConstantExtrapolationRule(), D.NullExtrapolationRule())
    # the output is the first npts rows of a larger array: point slots past
    # the end of a line must leave the guard rows after it untouched
    guard = 600
    full = torch.full((npts + guard, nbatch), np.nan, dtype=torch.float64, device=cuda)
    full[npts:] = -7.0
    out = full[:npts]
    ev(out, torch.from_numpy(x).to(cuda), torch.from_numpy(coef).to(cuda))
    assert bool((full[npts:] == -7.0).all())
    got = out.cpu().numpy()
    rng = np.random.default_rng(8)
    lines = np.unique(np.concatenate([np.arange(64), np.arange(nbatch - 64, nbatch),
                                      rng.choice(nbatch, 64, replace=False)]))
    ref = orc.eval(x[:, lines], coef[:, lines], extrap=("constant", "null"))
    assert not np.isnan(got).any()
    assert rel_err(got[:, lines], ref) <= TOL
    # deriv(1) takes the same persistent form for degrees up to 3 (no rule
    # applies: points inside the domain)
    xin = np.clip(x, orc.rmin, orc.rmax)
    full[:npts] = np.nan
    ev.deriv(1, out, torch.from_numpy(xin).to(cuda), torch.from_numpy(coef).to(cuda))
    assert bool((full[npts:] == -7.0).all())
    got = out.cpu().numpy()
    ref = orc.eval(xin[:, lines], coef[:, lines], deriv=1)
    assert not np.isnan(got).any()
    assert rel_err(got[:, lines], ref) <= TOL


@pytest.mark.parametrize("mesh,ncells,ratio", [("random", 300, 1.05), ("random", 1000, 1.05), ("jitter", 500, 1.05),
                                               ("geometric", 2048, 1.05), ("geometric", 700, 1.01)])
@pytest.mark.parametrize("layout", ["batch_contiguous", "interp_contiguous"])
def test_nonuniform_cell_search_at_edges(mesh, ncells, ratio, layout, cuda):
    """deriv(degree) is piecewise constant with a jump at every knot, so it
    exposes any wrong cell index: evaluate it at every break point, every
    search-bucket edge rmin + k (rmax - rmin) / ncells, and their floating-
    point neighbours, against the oracle's binary search."""
    import torch
    import ddc_amd as D

    case = Case(3, False, G, ncells=ncells, mesh=mesh, ratio=ratio, seed=11)
    orc, coef1 = _coef(case, 1, seed=3)
    bs, _ = case.product()
    br = orc.knots()[3: 3 + ncells + 1]
    edges = orc.rmin + np.arange(ncells + 1) * ((orc.rmax - orc.rmin) / ncells)
    base = np.concatenate([br, edges])
    x = np.concatenate([base, np.nextafter(base, np.inf), np.nextafter(base, -np.inf)])
    x = np.clip(x, orc.rmin, orc.rmax)
    nbatch = 8
    xs = np.repeat(x[:, None], nbatch, axis=1)
    coef = np.repeat(coef1, nbatch, axis=1)
    ref = orc.eval(xs, coef, deriv=3)
    ev = D.SplineEvaluator(bs, D.NullExtrapolationRule(), D.NullExtrapolationRule())
    if layout == "batch_contiguous":
        tx, tc = torch.from_numpy(xs).to(cuda), torch.from_numpy(coef).to(cuda)
        out = torch.empty_like(tx)
        ev.deriv(3, out, tx, tc, axis=0)
        got = out.cpu().numpy()
    else:
        tx = torch.from_numpy(np.ascontiguousarray(xs.T)).to(cuda)
        tc = torch.from_numpy(np.ascontiguousarray(coef.T)).to(cuda)
        out = torch.empty_like(tx)
        ev.deriv(3, out, tx, tc, axis=1)
        got = out.cpu().numpy().T
    assert rel_err(got, ref) <= TOL


@pytest.mark.parametrize("case", [Case(3, True, P, ncells=1024), Case(5, True, P, ncells=1024),
                                  Case(3, False, G, ncells=1024, mesh="random")], ids=lambda c: c.id)
@pytest.mark.parametrize("layout", ["batch_contiguous", "interp_contiguous"])
def test_deriv_orders_tight(case, layout, cuda):
    """deriv(k) on the FP64 workload's bases (uniform periodic cubic and
    quintic, non-uniform cubic Greville, 1024 cells), relative to max
    |deriv(k)| of the oracle: the first derivative at the 1e-13 bar of the
    values; every order's measured error is printed."""
    import torch
    import ddc_amd as D

    nbatch, npts = 64, 1024
    orc, coef = _coef(case, nbatch, seed=11)
    bs, _ = case.product()
    rng = np.random.default_rng(12)
    x = rng.uniform(orc.rmin, orc.rmax, (npts, nbatch))
    ev = D.SplineEvaluator(bs)
    bc = layout == "batch_contiguous"
    ax = 0 if bc else 1
    tc = torch.from_numpy(coef if bc else np.ascontiguousarray(coef.T)).to(cuda)
    tx = torch.from_numpy(x if bc else np.ascontiguousarray(x.T)).to(cuda)
    errs = []
    for order in range(1, case.degree + 1):
        ref = orc.eval(x, coef, deriv=order)
        out = torch.empty_like(tx)
        ev.deriv(order, out, tx, tc, axis=ax)
        got = out.cpu().numpy() if bc else out.cpu().numpy().T
        errs.append(rel_err(got, ref))
    print(f"{case.id} {layout}: deriv orders 1..{case.degree} rel err " + " ".join(f"{e:.1e}" for e in errs))
    assert errs[0] <= TOL
    for k, e in enumerate(errs, start=1):
        assert e <= TOL * 10 ** (k - 1), (k, e)


@pytest.mark.parametrize("case", [Case(3, True, P, ncells=1024), Case(3, False, G, ncells=500, mesh="random"),
                                  Case(5, False, H, ncells=300, mesh="random")], ids=lambda c: c.id)
@pytest.mark.parametrize("layout", ["batch_contiguous", "interp_contiguous"])
@pytest.mark.parametrize("nbatch", [37, 3], ids=["tiled", "per_point"])
def test_eval_at_domain_points(case, layout, nbatch, cuda):
    """operator()(spline_eval, spline_coef) / deriv(order, spline_eval,
    spline_coef): one coordinate vector shared by every line
    (ddcx_evaluator_eval_at), bit for bit against the same points repeated on
    every line, on the tiled kernels (37 lines) and the per-point kernel (3),
    points outside the domain for the value form, and against the oracle."""
    import torch
    import ddc_amd as D

    orc, coef = _coef(case, nbatch, seed=8)
    bs, _ = case.product()
    rng = np.random.default_rng(9)
    npts = 300
    w = orc.rmax - orc.rmin
    pts = rng.uniform(orc.rmin - 0.1 * w, orc.rmax + 0.1 * w, npts)
    ev = D.SplineEvaluator(bs, D.ConstantExtrapolationRule(), D.NullExtrapolationRule())
    bc = layout == "batch_contiguous"
    ax = 0 if bc else 1
    tc = torch.from_numpy(coef if bc else np.ascontiguousarray(coef.T)).to(cuda)
    for order in (0, 1):
        p = pts if order == 0 else np.clip(pts, orc.rmin, orc.rmax)
        full = np.repeat(p[:, None], nbatch, axis=1)
        tx = torch.from_numpy(full if bc else np.ascontiguousarray(full.T)).to(cuda)
        ref_out = torch.empty_like(tx)
        (ev(ref_out, tx, tc, axis=ax) if order == 0 else ev.deriv(order, ref_out, tx, tc, axis=ax))
        out = torch.full_like(tx, np.nan)
        ev.eval_at_points(out, torch.from_numpy(p).to(cuda), tc, axis=ax, order=order)
        assert np.array_equal(out.cpu().numpy(), ref_out.cpu().numpy())
        got = out.cpu().numpy() if bc else out.cpu().numpy().T
        ref = orc.eval(full, coef, deriv=order, extrap=("constant", "null"))
        assert np.abs(got - ref).max() / max(1.0, np.abs(ref).max()) <= TOL * (10 ** order)


@pytest.mark.parametrize("degree", [3, 5])
@pytest.mark.parametrize("scale", [1e-100, 1e-40, 1e100])
@pytest.mark.parametrize("layout", ["batch_contiguous", "interp_contiguous"])
def test_nonuniform_values_at_extreme_scales(degree, scale, layout, cuda):
    """A non-uniform basis inverts a point's d(d+1)/2 knot-difference divisors
    together, from one reciprocal of their product; a product outside
    [2^-960, 2^960] takes one reciprocal per divisor. Break points scaled to
    [0, 1e-100], [0, 1e-40] and [0, 1e100] send the cubic and quintic
    products below and above that range (1e-40 keeps the cubic one inside):
    both branches against the oracle's divisions."""
    import torch
    import ddc_amd as D
    from cases import breaks_for

    import oracle as O

    br = breaks_for("random", 200, seed=12) * scale
    orc = O.Spline(degree, False, G, G, breaks=br)
    bs = D.NonUniformBSplines(degree, br)
    rng = np.random.default_rng(13)
    nbatch, npts = 37, 400
    coef = rng.standard_normal((orc.size, nbatch))
    x = rng.uniform(orc.rmin, orc.rmax, (npts, nbatch))
    x[:201, 0] = br  # the break points themselves
    ref = orc.eval(x, coef)
    ev = D.SplineEvaluator(bs, D.NullExtrapolationRule(), D.NullExtrapolationRule())
    if layout == "batch_contiguous":
        tx, tc = torch.from_numpy(x).to(cuda), torch.from_numpy(coef).to(cuda)
        out = torch.empty_like(tx)
        ev(out, tx, tc, axis=0)
        got = out.cpu().numpy()
    else:
        tx = torch.from_numpy(np.ascontiguousarray(x.T)).to(cuda)
        tc = torch.from_numpy(np.ascontiguousarray(coef.T)).to(cuda)
        out = torch.empty_like(tx)
        ev(out, tx, tc, axis=1)
        got = out.cpu().numpy().T
    assert np.all(np.isfinite(got))
    assert rel_err(got, ref) <= TOL
