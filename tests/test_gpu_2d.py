"""GPU parity of the tensor-product path (SplineBuilder2D / SplineEvaluator2D,
spline_builder_2d.hpp:514-651, spline_evaluator_2d.hpp:370-825) against the
oracle's 2D restatement on the same inputs.

Bar: coefficients and evaluations <= 1e-13 relative (max norm).
"""
import numpy as np
import pytest

from cases import G, P, breaks_for
from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-13


def _pair(n1, n2, seed=0):
    """x: non-uniform cubic, Greville closures; y: uniform periodic cubic."""
    import ddc_amd as D
    import oracle as O

    br = breaks_for("random", n1, seed=seed)
    o1 = O.Spline(3, False, G, G, breaks=br)
    o2 = O.Spline(3, True, P, P, ncells=n2, rmin=0.0, rmax=1.0)
    bs1 = D.NonUniformBSplines(3, br)
    bs2 = D.UniformBSplines(3, 0.0, 1.0, n2, periodic=True)
    b1 = D.SplineBuilder(bs1, D.SplineBuilderClosure.GREVILLE, D.SplineBuilderClosure.GREVILLE)
    b2 = D.SplineBuilder(bs2)
    return o1, o2, bs1, bs2, b1, b2


def _vals(o1, o2, nb, seed):
    rng = np.random.default_rng(seed)
    x = o1.points()[None, :, None]
    y = o2.points()[None, None, :]
    k = rng.integers(1, 5, size=(nb, 1, 1))
    return np.sin(2 * np.pi * k * x) * np.cos(2 * np.pi * y) + 0.1 * rng.standard_normal((nb, x.shape[1], y.shape[2]))


@pytest.mark.parametrize("n1,n2,nb", [(20, 16, 3), (300, 256, 2)])
def test_build2d_matches_oracle(n1, n2, nb, cuda):
    import torch
    import ddc_amd as D
    import oracle as O

    o1, o2, bs1, bs2, b1, b2 = _pair(n1, n2)
    vals = _vals(o1, o2, nb, 1)
    ref = O.spline2d_build(o1, o2, vals)
    coef = torch.full((nb, bs1.size(), bs2.size()), np.nan, dtype=torch.float64, device=cuda)
    D.SplineBuilder2D(b1, b2)(coef, torch.from_numpy(vals).to(cuda))
    got = coef.cpu().numpy()
    assert np.all(np.isfinite(got))
    assert rel_err(got, ref) <= TOL


@pytest.mark.parametrize("n1,n2,nb", [(1024, 1024, 2), (300, 320, 3)])
def test_build2d_fp32_refine_matches_oracle(n1, n2, nb, cuda):
    """The tensor-product workload (non-uniform cubic Greville x random
    breakpoints x uniform periodic cubic, 1024 x 1024 at full size) built
    by SplineBuilder2D over two FP32-sweep + FP64-refinement builders,
    against the FP64 oracle: <= 1e-12 relative on every line of the final
    (dimension-2) pass and on every batch entry."""
    import torch
    import ddc_amd as D
    import oracle as O

    br = breaks_for("random", n1, seed=4)
    o1 = O.Spline(3, False, G, G, breaks=br)
    o2 = O.Spline(3, True, P, P, ncells=n2, rmin=0.0, rmax=1.0)
    bs1 = D.NonUniformBSplines(3, br)
    bs2 = D.UniformBSplines(3, 0.0, 1.0, n2, periodic=True)
    F = D.Precision.FP32_REFINE
    b1 = D.SplineBuilder(bs1, D.SplineBuilderClosure.GREVILLE, D.SplineBuilderClosure.GREVILLE, precision=F)
    b2 = D.SplineBuilder(bs2, precision=F)
    assert b1.info.kernel == 2 and b2.info.kernel == 2, "mixed-precision kernel not selected"
    vals = _vals(o1, o2, nb, 3)
    ref = O.spline2d_build(o1, o2, vals)
    coef = torch.full((nb, bs1.size(), bs2.size()), np.nan, dtype=torch.float64, device=cuda)
    D.SplineBuilder2D(b1, b2)(coef, torch.from_numpy(vals).to(cuda))
    got = coef.cpu().numpy()
    assert np.all(np.isfinite(got))
    line = (np.abs(got - ref).max(axis=2) / np.abs(ref).max(axis=2)).max()
    entry = max(rel_err(got[l], ref[l]) for l in range(nb))
    print(f"2D FP32+refine {n1}x{n2}: steps {b1.info.refine_steps}/{b2.info.refine_steps}, "
          f"max line err {line:.2e}, max entry err {entry:.2e}")
    assert line <= 1e-12 and entry <= 1e-12


@pytest.mark.parametrize("orders", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_eval2d_matches_oracle(orders, cuda):
    import torch
    import ddc_amd as D
    import oracle as O

    o1, o2, bs1, bs2, b1, b2 = _pair(24, 20, seed=2)
    nb, npts = 3, 500
    rng = np.random.default_rng(5)
    coef = rng.standard_normal((nb, bs1.size(), bs2.size()))
    coef[:, :, bs2.nbasis():] = coef[:, :, :3]  # periodic duplicates along y
    coords = np.stack([rng.uniform(0, 1, (nb, npts)), rng.uniform(-0.5, 1.5, (nb, npts))], axis=-1)
    ref = O.spline2d_eval(o1, o2, coords, coef, *orders)
    ev1 = D.SplineEvaluator(bs1, D.NullExtrapolationRule(), D.NullExtrapolationRule())
    ev2 = D.SplineEvaluator(bs2, D.PeriodicExtrapolationRule(), D.PeriodicExtrapolationRule())
    out = torch.full((nb, npts), np.nan, dtype=torch.float64, device=cuda)
    D.SplineEvaluator2D(ev1, ev2)(out, torch.from_numpy(coords).to(cuda), torch.from_numpy(coef).to(cuda), *orders)
    assert rel_err(out.cpu().numpy(), ref) <= TOL


def test_build_eval2d_reproduces_interpolation_points(cuda):
    """Evaluating the 2D spline at the interpolation grid returns the values."""
    import torch
    import ddc_amd as D

    o1, o2, bs1, bs2, b1, b2 = _pair(40, 32, seed=4)
    vals = _vals(o1, o2, 2, 3)
    coef = torch.empty((2, bs1.size(), bs2.size()), dtype=torch.float64, device=cuda)
    D.SplineBuilder2D(b1, b2)(coef, torch.from_numpy(vals).to(cuda))
    X, Y = np.meshgrid(o1.points(), o2.points(), indexing="ij")
    coords = np.broadcast_to(np.stack([X.ravel(), Y.ravel()], -1), (2, X.size, 2)).copy()
    ev1 = D.SplineEvaluator(bs1, D.NullExtrapolationRule(), D.NullExtrapolationRule())
    ev2 = D.SplineEvaluator(bs2, D.PeriodicExtrapolationRule(), D.PeriodicExtrapolationRule())
    out = torch.empty((2, X.size), dtype=torch.float64, device=cuda)
    D.SplineEvaluator2D(ev1, ev2)(out, torch.from_numpy(coords).to(cuda), coef)
    assert rel_err(out.cpu().numpy(), vals.reshape(2, -1)) <= 1e-12


def _small_pair():
    return _pair(8, 8, seed=6)


def test_eval2d_vjp_jacobian_and_fd(cuda):
    """gcoef against the Jacobian of the oracle's 2D evaluation in the
    coefficients (unit tensors); gx against w * oracle partial derivatives and
    central differences (h = 1e-6)."""
    import torch
    import ddc_amd as D
    import oracle as O

    o1, o2, bs1, bs2, b1, b2 = _small_pair()
    s1, s2 = bs1.size(), bs2.size()
    nb, npts = 2, 64
    rng = np.random.default_rng(8)
    coef = rng.standard_normal((nb, s1, s2))
    coords = np.stack([rng.uniform(0.01, 0.99, (nb, npts)), rng.uniform(-0.3, 1.3, (nb, npts))], axis=-1)
    w = rng.standard_normal((nb, npts))
    ev1 = D.SplineEvaluator(bs1, D.NullExtrapolationRule(), D.NullExtrapolationRule())
    ev2 = D.SplineEvaluator(bs2, D.PeriodicExtrapolationRule(), D.PeriodicExtrapolationRule())
    ev = D.SplineEvaluator2D(ev1, ev2)
    tc = torch.from_numpy(coef).to(cuda)
    gx = torch.full((nb, npts, 2), np.nan, dtype=torch.float64, device=cuda)
    gc = torch.full_like(tc, np.nan)
    ev.eval_vjp(torch.from_numpy(coords).to(cuda), torch.from_numpy(w).to(cuda), tc, gx, gc)
    gx, gc = gx.cpu().numpy(), gc.cpu().numpy()
    # Jacobian in the coefficients: all unit tensors as one oracle batch per line
    units = np.eye(s1 * s2).reshape(s1 * s2, s1, s2)
    for l in range(nb):
        J = O.spline2d_eval(o1, o2, np.broadcast_to(coords[l], (s1 * s2, npts, 2)).copy(), units)  # (s1*s2, npts)
        ref = (J @ w[l]).reshape(s1, s2)
        assert np.abs(gc[l] - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())
    # deriv_dim_* evaluate without boundary handling (eval_no_bc), so the
    # reference partials are taken at the periodically wrapped coordinates
    wrapped = coords.copy()
    wrapped[..., 1] -= np.floor(wrapped[..., 1])
    d1 = O.spline2d_eval(o1, o2, wrapped, coef, 1, 0)
    d2 = O.spline2d_eval(o1, o2, wrapped, coef, 0, 1)
    assert rel_err(gx[..., 0], w * d1) <= 1e-13
    assert rel_err(gx[..., 1], w * d2) <= 1e-13
    h = 1e-6
    for k in range(2):
        cp, cm = coords.copy(), coords.copy()
        cp[..., k] += h
        cm[..., k] -= h
        fd = (O.spline2d_eval(o1, o2, cp, coef) - O.spline2d_eval(o1, o2, cm, coef)) / (2 * h)
        assert np.abs(gx[..., k] - w * fd).max() <= 1e-6 * max(1.0, np.abs(w * fd).max())


def test_build2d_vjp_is_jacobian_transpose(cuda):
    import torch
    import ddc_amd as D
    import oracle as O

    o1, o2, bs1, bs2, b1, b2 = _small_pair()
    n1, n2 = o1.npts, o2.npts
    s1, s2 = bs1.size(), bs2.size()
    units = np.eye(n1 * n2).reshape(n1 * n2, n1, n2)
    J = O.spline2d_build(o1, o2, units).reshape(n1 * n2, s1 * s2)  # row k: build(e_k)
    rng = np.random.default_rng(9)
    g = rng.standard_normal((3, s1, s2))
    ref = (g.reshape(3, -1) @ J.T).reshape(3, n1, n2)
    gv = torch.full((3, n1, n2), np.nan, dtype=torch.float64, device=cuda)
    D.SplineBuilder2D(b1, b2).build_vjp(torch.from_numpy(g).to(cuda), gv)
    assert rel_err(gv.cpu().numpy(), ref) <= 1e-13


HERMITE_2D = [
    # (degree, uniform, lower, upper, ncells) per dimension
    ((3, False, "H", "H", 24), (5, True, "H", "H", 20)),
    ((3, True, "P", "P", 16), (3, False, "H", "G", 22)),
    ((4, True, "H", "HH", 18), (3, True, "G", "H", 15)),
    ((5, False, "HH", "H", 21), (2, True, "P", "P", 12)),
]


@pytest.mark.parametrize("dims", HERMITE_2D, ids=["HHxHH", "PxHG", "HHHxGH", "HHHxP"])
def test_build2d_hermite_matches_oracle(dims, cuda):
    """SplineBuilder2D with Hermite closures (spline_builder_2d.hpp:514-651):
    boundary derivatives of each dimension and the corner cross derivatives,
    against the oracle's restatement."""
    import torch
    import ddc_amd as D
    import oracle as O

    bc = {"H": O.HERMITE, "HH": O.HOMOGENEOUS_HERMITE, "G": G, "P": P}
    orcs, builders = [], []
    for i, (d, uni, lo, up, n) in enumerate(dims):
        br = None if uni else breaks_for("random", n, seed=20 + i)
        periodic = lo == "P"
        if uni:
            o = O.Spline(d, True, bc[lo], bc[up], ncells=n, rmin=0.0, rmax=1.0)
            bs = D.UniformBSplines(d, 0.0, 1.0, n, periodic=periodic)
        else:
            o = O.Spline(d, False, bc[lo], bc[up], breaks=br)
            bs = D.NonUniformBSplines(d, br, periodic=periodic)
        orcs.append(o)
        builders.append(D.SplineBuilder(bs, D.SplineBuilderClosure(bc[lo]), D.SplineBuilderClosure(bc[up])))
    o1, o2 = orcs
    nb = 3
    rng = np.random.default_rng(9)
    r = lambda *s: rng.standard_normal(s)  # noqa: E731
    vals = r(nb, o1.npts, o2.npts)
    H1l, H1u = dims[0][2] == "H", dims[0][3] == "H"
    H2l, H2u = dims[1][2] == "H", dims[1][3] == "H"
    dmin1 = r(nb, o1.nbv_l, o2.npts) if H1l else None
    dmax1 = r(nb, o1.nbv_r, o2.npts) if H1u else None
    dmin2 = r(nb, o1.npts, o2.nbv_l) if H2l else None
    dmax2 = r(nb, o1.npts, o2.nbv_r) if H2u else None
    mm = r(nb, o1.nbv_l, o2.nbv_l) if H1l and H2l else None
    Mm = r(nb, o1.nbv_r, o2.nbv_l) if H1u and H2l else None
    mM = r(nb, o1.nbv_l, o2.nbv_r) if H1l and H2u else None
    MM = r(nb, o1.nbv_r, o2.nbv_r) if H1u and H2u else None
    ref = O.spline2d_build(o1, o2, vals, dmin1, dmax1, dmin2, dmax2, mm, Mm, mM, MM)
    t = lambda a: None if a is None else torch.from_numpy(a).to(cuda)  # noqa: E731
    coef = torch.full((nb, o1.size, o2.size), np.nan, dtype=torch.float64, device=cuda)
    D.SplineBuilder2D(*builders)(coef, t(vals), t(dmin1), t(dmax1), t(dmin2), t(dmax2), t(mm), t(Mm), t(mM), t(MM))
    got = coef.cpu().numpy()
    assert np.all(np.isfinite(got))
    assert rel_err(got, ref) <= TOL


@pytest.mark.parametrize("dims", HERMITE_2D, ids=["HHxHH", "PxHG", "HHHxGH", "HHHxP"])
def test_build2d_hermite_vjp_is_jacobian_transpose(dims, cuda):
    """SplineBuilder2D.build_vjp with Hermite closures: the gradient of every
    input block (values, the boundary derivatives of each dimension, the
    corner cross derivatives) equals the transposed Jacobian of the oracle's
    2D build for that block, assembled from unit inputs with the other
    blocks zero.  Bars: 1e-13 for the values, 1e-12 for the derivative
    blocks, whose gradients pass through the Hermite border rows."""
    import torch
    import ddc_amd as D
    import oracle as O

    bc = {"H": O.HERMITE, "HH": O.HOMOGENEOUS_HERMITE, "G": G, "P": P}
    orcs, builders = [], []
    for i, (d, uni, lo, up, n) in enumerate(dims):
        br = None if uni else breaks_for("random", n, seed=20 + i)
        periodic = lo == "P"
        if uni:
            o = O.Spline(d, True, bc[lo], bc[up], ncells=n, rmin=0.0, rmax=1.0)
            bs = D.UniformBSplines(d, 0.0, 1.0, n, periodic=periodic)
        else:
            o = O.Spline(d, False, bc[lo], bc[up], breaks=br)
            bs = D.NonUniformBSplines(d, br, periodic=periodic)
        orcs.append(o)
        builders.append(D.SplineBuilder(bs, D.SplineBuilderClosure(bc[lo]), D.SplineBuilderClosure(bc[up])))
    o1, o2 = orcs
    H1l, H1u = dims[0][2] == "H", dims[0][3] == "H"
    H2l, H2u = dims[1][2] == "H", dims[1][3] == "H"
    n1, n2 = o1.npts, o2.npts
    # block name -> shape of one batch entry (None: the closures take no such input)
    shapes = {"vals": (n1, n2),
              "dmin1": (o1.nbv_l, n2) if H1l else None, "dmax1": (o1.nbv_r, n2) if H1u else None,
              "dmin2": (n1, o2.nbv_l) if H2l else None, "dmax2": (n1, o2.nbv_r) if H2u else None,
              "mm": (o1.nbv_l, o2.nbv_l) if H1l and H2l else None, "Mm": (o1.nbv_r, o2.nbv_l) if H1u and H2l else None,
              "mM": (o1.nbv_l, o2.nbv_r) if H1l and H2u else None, "MM": (o1.nbv_r, o2.nbv_r) if H1u and H2u else None}
    order = ["vals", "dmin1", "dmax1", "dmin2", "dmax2", "mm", "Mm", "mM", "MM"]
    nb = 2
    gs = np.random.default_rng(11).standard_normal((nb, o1.size, o2.size))
    outs = {k: torch.full((nb,) + sh, np.nan, dtype=torch.float64, device=cuda)
            for k, sh in shapes.items() if sh is not None}
    D.SplineBuilder2D(*builders).build_vjp(torch.from_numpy(gs).to(cuda), *[outs.get(k) for k in order])
    for k, sh in shapes.items():
        if sh is None:
            continue
        m = int(np.prod(sh))
        args = {j: (np.zeros((m,) + s2) if s2 is not None else None) for j, s2 in shapes.items()}
        args[k] = np.eye(m).reshape((m,) + sh)
        J = O.spline2d_build(o1, o2, *[args[j] for j in order]).reshape(m, -1)  # row e: build(unit e of block k)
        ref = gs.reshape(nb, -1) @ J.T
        got = outs[k].cpu().numpy().reshape(nb, m)
        assert np.all(np.isfinite(got)), k
        assert rel_err(got, ref) <= (TOL if k == "vals" else 1e-12), k


@pytest.mark.parametrize("rules", [("null", "periodic"), ("constant", "periodic"), ("constant", "constant")])
def test_eval2d_vjp_binned_matches_atomic(rules, cuda):
    """The binned 2D adjoint (ddcx_evaluator2d_eval_vjp_binned: points
    grouped by cell tile, LDS accumulation) against the one-atomic-per-term
    kernel (ddcx_evaluator2d_eval_vjp), with points outside the domain on
    every side; 16 batch entries, ragged tile occupancy."""
    import ctypes as C
    import torch
    import ddc_amd as D
    from ddc_amd._lib import call

    br = breaks_for("random", 70, seed=5)
    bs1 = D.NonUniformBSplines(3, br)
    per2 = rules[1] == "periodic"
    bs2 = D.UniformBSplines(2, 0.0, 1.0, 45, periodic=per2)
    mk = {"null": D.NullExtrapolationRule, "constant": D.ConstantExtrapolationRule,
          "periodic": D.PeriodicExtrapolationRule}
    ev = D.SplineEvaluator2D(D.SplineEvaluator(bs1, mk[rules[0]](), mk[rules[0]]()),
                             D.SplineEvaluator(bs2, mk[rules[1]](), mk[rules[1]]()))
    nb, npts = 16, 3000
    rng = np.random.default_rng(11)
    coef = torch.from_numpy(rng.standard_normal((nb, bs1.size(), bs2.size()))).to(cuda)
    xy = torch.from_numpy(np.stack([rng.uniform(-0.2, 1.2, (nb, npts)), rng.uniform(-0.3, 1.3, (nb, npts))],
                                   axis=-1)).to(cuda)
    w = torch.from_numpy(rng.standard_normal((nb, npts))).to(cuda)
    need = C.c_int64(0)
    call("ddcx_evaluator2d_eval_vjp_workspace", ev.ev1._h, ev.ev2._h, npts, nb, C.byref(need))
    assert need.value > 0
    gx_b, gc_b = torch.full_like(xy, np.nan), torch.full_like(coef, np.nan)
    ev.eval_vjp(xy, w, coef, gx_b, gc_b)  # binned
    gx_a, gc_a = torch.full_like(xy, np.nan), torch.full_like(coef, np.nan)
    call("ddcx_evaluator2d_eval_vjp", ev.ev1._h, ev.ev2._h, xy.data_ptr(), w.data_ptr(), gx_a.data_ptr(), npts,
         coef.data_ptr(), gc_a.data_ptr(), coef.stride(0), coef.stride(1), coef.stride(2), nb, None)
    torch.cuda.synchronize()
    gb, ga = gc_b.cpu().numpy(), gc_a.cpu().numpy()
    assert np.all(np.isfinite(gb)) and np.all(np.isfinite(gx_b.cpu().numpy()))
    assert np.abs(gb - ga).max() <= 1e-12 * np.abs(ga).max()
    assert np.array_equal(gx_b.cpu().numpy(), gx_a.cpu().numpy())
    # a smaller call reuses the workspace
    ev.eval_vjp(xy[:, :100].contiguous(), w[:, :100].contiguous(), coef, gx_b[:, :100].contiguous(), gc_b)
    assert np.all(np.isfinite(gc_b.cpu().numpy()))


@pytest.mark.parametrize("degs,orders", [((3, 3), (0, 0)), ((3, 3), (1, 1)), ((2, 5), (0, 2)), ((5, 4), (2, 0)),
                                         ((1, 2), (0, 0))])
def test_eval2d_packed_matches_oracle(degs, orders, cuda):
    """The packed-coefficient evaluation (ddcx_evaluator2d_eval_packed):
    values and partial derivatives against the oracle, with points outside
    the domain on every side (Null x periodic); a workspace of one batch
    entry (entries packed and evaluated one at a time) gives the same
    values as a workspace of all."""
    import ctypes as C
    import torch
    import ddc_amd as D
    import oracle as O
    from ddc_amd._lib import call

    d1, d2 = degs
    br = breaks_for("random", 23, seed=9)
    o1 = O.Spline(d1, False, G, G, breaks=br)
    o2 = O.Spline(d2, True, P, P, ncells=17, rmin=0.0, rmax=1.0)
    bs1 = D.NonUniformBSplines(d1, br)
    bs2 = D.UniformBSplines(d2, 0.0, 1.0, 17, periodic=True)
    ev = D.SplineEvaluator2D(D.SplineEvaluator(bs1, D.NullExtrapolationRule(), D.NullExtrapolationRule()),
                             D.SplineEvaluator(bs2, D.PeriodicExtrapolationRule(), D.PeriodicExtrapolationRule()))
    nb, npts = 3, 2000
    rng = np.random.default_rng(13)
    coef = rng.standard_normal((nb, bs1.size(), bs2.size()))
    coef[:, :, bs2.nbasis():] = coef[:, :, :d2]
    lo = 0.0 if orders != (0, 0) else -0.3
    coords = np.stack([rng.uniform(lo, 1 - lo, (nb, npts)), rng.uniform(-0.5, 1.5, (nb, npts))], axis=-1)
    if orders != (0, 0):  # derivatives apply no wrap: keep dimension 2 inside
        coords[..., 1] = rng.uniform(0.0, 1.0, (nb, npts))
    ref = O.spline2d_eval(o1, o2, coords, coef, *orders)
    tc, tx = torch.from_numpy(coef).to(cuda), torch.from_numpy(coords).to(cuda)
    out = torch.full((nb, npts), np.nan, dtype=torch.float64, device=cuda)
    ev(out, tx, tc, *orders, packed=True)
    assert rel_err(out.cpu().numpy(), ref) <= TOL
    per = C.c_int64(0)
    call("ddcx_evaluator2d_eval_workspace", ev.ev1._h, ev.ev2._h, 1, C.byref(per))
    ws = torch.empty(per.value + 100, dtype=torch.uint8, device=cuda)  # one entry: packed one at a time
    out1 = torch.full_like(out, np.nan)
    call("ddcx_evaluator2d_eval_packed", ev.ev1._h, ev.ev2._h, orders[0], orders[1], out1.data_ptr(), tx.data_ptr(),
         npts, tc.data_ptr(), tc.stride(0), tc.stride(1), tc.stride(2), nb, ws.data_ptr(), ws.numel(), None)
    torch.cuda.synchronize()
    assert np.array_equal(out1.cpu().numpy(), out.cpu().numpy())


@pytest.mark.parametrize("n1,n2", [(23, 19), (6000, 21)], ids=["both_staged", "dim1_too_large_for_lds"])
def test_eval2d_packed_nonuniform_both(n1, n2, cuda):
    """Packed evaluation with both dimensions non-uniform: the knots and cell
    brackets of both staged in LDS, and of dimension 1 read from global
    memory when they do not fit (6000 cells); constant rules on dimension 2
    with points outside on every side, against the oracle."""
    import torch
    import ddc_amd as D
    import oracle as O

    br1 = breaks_for("random", n1, seed=21)
    br2 = breaks_for("random", n2, seed=22)
    o1 = O.Spline(3, False, G, G, breaks=br1)
    o2 = O.Spline(2, False, G, G, breaks=br2)
    bs1, bs2 = D.NonUniformBSplines(3, br1), D.NonUniformBSplines(2, br2)
    ev = D.SplineEvaluator2D(D.SplineEvaluator(bs1, D.ConstantExtrapolationRule(), D.ConstantExtrapolationRule()),
                             D.SplineEvaluator(bs2, D.ConstantExtrapolationRule(), D.ConstantExtrapolationRule()))
    nb, npts = 2, 3000
    rng = np.random.default_rng(23)
    coef = rng.standard_normal((nb, bs1.size(), bs2.size()))
    lo1, hi1, lo2, hi2 = br1[0], br1[-1], br2[0], br2[-1]
    w1, w2 = hi1 - lo1, hi2 - lo2
    coords = np.stack([rng.uniform(lo1 - 0.2 * w1, hi1 + 0.2 * w1, (nb, npts)),
                       rng.uniform(lo2 - 0.2 * w2, hi2 + 0.2 * w2, (nb, npts))], axis=-1)
    ref = O.spline2d_eval(o1, o2, coords, coef, 0, 0, rules=(1, 1, 1, 1))
    out = torch.full((nb, npts), np.nan, dtype=torch.float64, device=cuda)
    ev(out, torch.from_numpy(coords).to(cuda), torch.from_numpy(coef).to(cuda), packed=True)
    assert rel_err(out.cpu().numpy(), ref) <= TOL


@pytest.mark.parametrize("npts", [1, 37, 64, 257])
def test_eval2d_packed_ragged(npts, cuda):
    """The packed evaluation with fewer points than a wave or a workgroup, or
    one more than a workgroup's 256: every point against the oracle, with
    points outside the domain (Null x periodic)."""
    import torch
    import ddc_amd as D
    import oracle as O

    o1, o2, bs1, bs2, _, _ = _pair(12, 9, seed=3)
    ev = D.SplineEvaluator2D(D.SplineEvaluator(bs1, D.NullExtrapolationRule(), D.NullExtrapolationRule()),
                             D.SplineEvaluator(bs2, D.PeriodicExtrapolationRule(), D.PeriodicExtrapolationRule()))
    nb = 5
    rng = np.random.default_rng(npts)
    coef = rng.standard_normal((nb, bs1.size(), bs2.size()))
    coef[:, :, bs2.nbasis():] = coef[:, :, :3]
    coords = np.stack([rng.uniform(-0.2, 1.2, (nb, npts)), rng.uniform(-0.5, 1.5, (nb, npts))], axis=-1)
    ref = O.spline2d_eval(o1, o2, coords, coef)
    out = torch.full((nb, npts), np.nan, dtype=torch.float64, device=cuda)
    ev(out, torch.from_numpy(coords).to(cuda), torch.from_numpy(coef).to(cuda), packed=True)
    assert rel_err(out.cpu().numpy(), ref) <= TOL
