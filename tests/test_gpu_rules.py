"""GPU parity of the extrapolation rules with their constructor arguments:
ConstantExtrapolationRule<DimI>(eval_pos) in 1D and
ConstantExtrapolationRule<DimI, DimNI>(eval_pos, min, max) in 2D
(constant_extrapolation_rule.hpp:27-195), through
ddcx_evaluator_create_with_rules, on every kernel that applies a rule: the
per-point, LDS-tiled and persistent evaluators, both adjoints, the fused
build -> eval kernel, and the 2D evaluator and its two adjoints.  The oracle
restates the rules (pinned in test_oracle_eval.py against scipy at the rule
positions).  Bars: values and coordinate gradients <= 1e-13, coefficient
gradients <= 1e-12 (relative, max norm)."""
import numpy as np
import pytest

from cases import Case, G, H, P, breaks_for
from conftest import rel_err

pytestmark = pytest.mark.gpu


def _setup(case, nbatch, npts, seed, outside=0.3):
    import ddc_amd as D

    orc = case.oracle()
    bs, _ = case.product()
    rng = np.random.default_rng(seed)
    coef = rng.standard_normal((orc.size, nbatch))
    L = orc.rmax - orc.rmin
    x = rng.uniform(orc.rmin - outside * L, orc.rmax + outside * L, (npts, nbatch))
    pos = (orc.rmin + 0.37 * L, orc.rmin + 0.81 * L)
    ev = D.SplineEvaluator(bs, D.ConstantExtrapolationRule(pos[0]), D.ConstantExtrapolationRule(pos[1]))
    return orc, bs, coef, x, pos, ev, rng


CASES_1D = [Case(3, True, G, ncells=300), Case(3, False, G, ncells=257, mesh="random"),
            Case(5, False, H, ncells=400, mesh="random"), Case(2, True, G, ncells=64)]


@pytest.mark.parametrize("case", CASES_1D, ids=lambda c: c.id)
@pytest.mark.parametrize("shape", [(20, 3), (300, 37), (500, 8200)], ids=["point", "tiled", "stream"])
def test_eval_constant_eval_pos(case, shape, cuda):
    """Per-point (few lines), LDS-tiled and, for uniform bases on 8200
    batch-contiguous lines, the persistent streaming evaluator."""
    import torch

    npts, nbatch = shape
    orc, bs, coef, x, pos, ev, _ = _setup(case, nbatch, npts, 1)
    ref = orc.eval(x, coef, extrap=("constant", "constant"), eval_pos=pos)
    bound = orc.eval(x, coef, extrap=("constant", "constant"))
    assert rel_err(bound, ref) > 1e-6  # the positions matter on these points
    out = torch.full((npts, nbatch), np.nan, dtype=torch.float64, device=cuda)
    ev(out, torch.from_numpy(x).to(cuda), torch.from_numpy(coef).to(cuda))
    assert rel_err(out.cpu().numpy(), ref) <= 1e-13
    if nbatch <= 64:  # interpolation-contiguous form
        out = torch.full((nbatch, npts), np.nan, dtype=torch.float64, device=cuda)
        ev(out, torch.from_numpy(np.ascontiguousarray(x.T)).to(cuda),
           torch.from_numpy(np.ascontiguousarray(coef.T)).to(cuda), axis=1)
        assert rel_err(out.cpu().numpy().T, ref) <= 1e-13


@pytest.mark.parametrize("case", CASES_1D, ids=lambda c: c.id)
@pytest.mark.parametrize("shape", [(20, 3), (300, 21)], ids=["point", "tiled"])
def test_eval_vjp_constant_eval_pos(case, shape, cuda):
    """Both adjoint kernels: basis weights at eval_pos, zero coordinate
    gradient outside."""
    import torch

    npts, nbatch = shape
    orc, bs, coef, x, pos, ev, rng = _setup(case, nbatch, npts, 2)
    w = rng.standard_normal((npts, nbatch))
    gc_ref, gx_ref = orc.eval_vjp(x, w, coef, extrap=("constant", "constant"), eval_pos=pos)
    tx, tw, tc = (torch.from_numpy(a).to(cuda) for a in (x, w, coef))
    gx, gc = torch.full_like(tx, np.nan), torch.full_like(tc, np.nan)
    ev.eval_vjp(tx, tw, tc, gx, gc)
    assert rel_err(gx.cpu().numpy(), gx_ref) <= 1e-13
    assert rel_err(gc.cpu().numpy(), gc_ref) <= 1e-12


@pytest.mark.parametrize("case", [Case(3, True, G, ncells=300), Case(5, False, G, ncells=400, mesh="random")],
                         ids=lambda c: c.id)
def test_build_eval_constant_eval_pos(case, cuda):
    """The fused build -> eval kernel applies the same rule positions."""
    import torch
    import ddc_amd as D

    nbatch, npts = 64, 400
    orc, bs, _, x, pos, _, rng = _setup(case, nbatch, npts, 3)
    bs, builder = case.product()
    ev = D.SplineEvaluator(bs, D.ConstantExtrapolationRule(pos[0]), D.ConstantExtrapolationRule(pos[1]))
    vals = rng.standard_normal((orc.npts, nbatch))
    coef_ref = orc.build(vals)
    ref = orc.eval(x, coef_ref, extrap=("constant", "constant"), eval_pos=pos)
    out = torch.full((npts, nbatch), np.nan, dtype=torch.float64, device=cuda)
    D.build_eval(builder, ev, out, torch.from_numpy(x).to(cuda), torch.from_numpy(vals).to(cuda))
    assert rel_err(out.cpu().numpy(), ref) <= 1e-13


def test_eval_pos_outside_domain_is_refused(cuda):
    import ddc_amd as D
    from ddc_amd._lib import DDCXError

    bs = D.UniformBSplines(3, 0.0, 1.0, 10)
    with pytest.raises(DDCXError, match="inside the domain"):
        D.SplineEvaluator(bs, D.ConstantExtrapolationRule(1.5), D.NullExtrapolationRule())
    with pytest.raises(ValueError):
        D.ConstantExtrapolationRule(0.5, 0.0)


def _pair2d(dim2_periodic):
    import ddc_amd as D
    import oracle as O

    br = breaks_for("random", 60, seed=3)
    o1 = O.Spline(3, False, G, G, breaks=br)
    bs1 = D.NonUniformBSplines(3, br)
    if dim2_periodic:
        o2 = O.Spline(2, True, P, P, ncells=45, rmin=0.0, rmax=1.0)
        bs2 = D.UniformBSplines(2, 0.0, 1.0, 45, periodic=True)
    else:
        o2 = O.Spline(2, True, G, G, ncells=45, rmin=-1.0, rmax=2.0)
        bs2 = D.UniformBSplines(2, -1.0, 2.0, 45)
    return o1, o2, bs1, bs2


def _rules2d(o1, o2, bs1, bs2, dim2_periodic, custom):
    """The reference's 2D constant rules (extrapolation_rule.cpp:238-256
    builds them at the domain bounds; `custom` moves them inside)."""
    import ddc_amd as D

    L1, L2 = o1.rmax - o1.rmin, o2.rmax - o2.rmin
    if custom:
        pos = [o1.rmin + 0.2 * L1, o1.rmin + 0.9 * L1, o2.rmin + 0.35 * L2, o2.rmin + 0.6 * L2]
        lo = [o2.rmin + 0.1 * L2, o2.rmin + 0.15 * L2, o1.rmin + 0.25 * L1, o1.rmin + 0.05 * L1]
        hi = [o2.rmin + 0.8 * L2, o2.rmin + 0.7 * L2, o1.rmin + 0.75 * L1, o1.rmin + 0.95 * L1]
    else:
        pos = [o1.rmin, o1.rmax, o2.rmin, o2.rmax]
        lo = [o2.rmin, o2.rmin, o1.rmin, o1.rmin]
        hi = [o2.rmax, o2.rmax, o1.rmax, o1.rmax]
    C = D.ConstantExtrapolationRule
    if dim2_periodic:
        ev1 = D.SplineEvaluator(bs1, C(pos[0]), C(pos[1]))
        ev2 = D.SplineEvaluator(bs2, D.PeriodicExtrapolationRule(), D.PeriodicExtrapolationRule())
        rules = (1, 1, 0, 0)
    else:
        ev1 = D.SplineEvaluator(bs1, C(pos[0], lo[0], hi[0]), C(pos[1], lo[1], hi[1]))
        ev2 = D.SplineEvaluator(bs2, C(pos[2], lo[2], hi[2]), C(pos[3], lo[3], hi[3]))
        rules = (1, 1, 1, 1)
    return D.SplineEvaluator2D(ev1, ev2), rules, (pos, lo, hi)


@pytest.mark.parametrize("dim2_periodic", [False, True], ids=["constant_constant", "constant_periodic"])
@pytest.mark.parametrize("custom", [False, True], ids=["bounds", "eval_pos"])
def test_eval2d_constant_rules(dim2_periodic, custom, cuda):
    """2D values with points outside on all four sides and at the corners,
    against the oracle's SplineEvaluator2D restatement."""
    import torch
    import oracle as O

    o1, o2, bs1, bs2 = _pair2d(dim2_periodic)
    ev, rules, rp = _rules2d(o1, o2, bs1, bs2, dim2_periodic, custom)
    nb, npts = 4, 3000
    rng = np.random.default_rng(12)
    coef = rng.standard_normal((nb, bs1.size(), bs2.size()))
    if dim2_periodic:
        coef[:, :, bs2.nbasis():] = coef[:, :, : bs2.degree()]
    L1, L2 = o1.rmax - o1.rmin, o2.rmax - o2.rmin
    coords = np.stack([rng.uniform(o1.rmin - 0.4 * L1, o1.rmax + 0.4 * L1, (nb, npts)),
                       rng.uniform(o2.rmin - 0.4 * L2, o2.rmax + 0.4 * L2, (nb, npts))], axis=-1)
    ref = O.spline2d_eval(o1, o2, coords, coef, rules=rules, rule_pos=rp)
    out = torch.full((nb, npts), np.nan, dtype=torch.float64, device=cuda)
    ev(out, torch.from_numpy(coords).to(cuda), torch.from_numpy(coef).to(cuda))
    assert rel_err(out.cpu().numpy(), ref) <= 1e-13


@pytest.mark.parametrize("dim2_periodic", [False, True], ids=["constant_constant", "constant_periodic"])
def test_eval2d_vjp_constant_eval_pos(dim2_periodic, cuda):
    """Both 2D adjoints with moved rule positions: gcoef against the
    Jacobian of the oracle's 2D evaluation in the coefficients, gx against
    central differences of the oracle (h = 1e-6) away from the places where
    the rules switch."""
    import ctypes as C
    import torch
    import oracle as O
    from ddc_amd._lib import call

    o1, o2, bs1, bs2 = _pair2d(dim2_periodic)
    ev, rules, rp = _rules2d(o1, o2, bs1, bs2, dim2_periodic, True)
    s1, s2 = bs1.size(), bs2.size()
    nb, npts = 2, 400
    rng = np.random.default_rng(13)
    coef = rng.standard_normal((nb, s1, s2))
    L1, L2 = o1.rmax - o1.rmin, o2.rmax - o2.rmin
    coords = np.stack([rng.uniform(o1.rmin - 0.4 * L1, o1.rmax + 0.4 * L1, (nb, npts)),
                       rng.uniform(o2.rmin - 0.4 * L2, o2.rmax + 0.4 * L2, (nb, npts))], axis=-1)
    w = rng.standard_normal((nb, npts))
    tc, txy, tw = (torch.from_numpy(a).to(cuda) for a in (coef, coords, w))
    gx_b, gc_b = torch.full_like(txy, np.nan), torch.full_like(tc, np.nan)
    ev.eval_vjp(txy, tw, tc, gx_b, gc_b)  # binned when the workspace query allows
    gx_a, gc_a = torch.full_like(txy, np.nan), torch.full_like(tc, np.nan)
    call("ddcx_evaluator2d_eval_vjp", ev.ev1._h, ev.ev2._h, txy.data_ptr(), tw.data_ptr(), gx_a.data_ptr(), npts,
         tc.data_ptr(), gc_a.data_ptr(), tc.stride(0), tc.stride(1), tc.stride(2), nb, None)
    torch.cuda.synchronize()
    units = np.eye(s1 * s2).reshape(s1 * s2, s1, s2)
    for gc in (gc_a.cpu().numpy(), gc_b.cpu().numpy()):
        for l in range(nb):
            J = O.spline2d_eval(o1, o2, np.broadcast_to(coords[l], (s1 * s2, npts, 2)).copy(), units, rules=rules,
                                rule_pos=rp)
            assert rel_err(gc[l], (J @ w[l]).reshape(s1, s2)) <= 1e-12
    h = 1e-6
    for gx in (gx_a.cpu().numpy(), gx_b.cpu().numpy()):
        for k in range(2):
            cp, cm = coords.copy(), coords.copy()
            cp[..., k] += h
            cm[..., k] -= h
            fp = O.spline2d_eval(o1, o2, cp, coef, rules=rules, rule_pos=rp)
            fm = O.spline2d_eval(o1, o2, cm, coef, rules=rules, rule_pos=rp)
            fd = w * (fp - fm) / (2 * h)
            # points within 2h of a domain bound or a clamp bound switch rule
            # inside the stencil; central differences do not apply there
            edges1 = [o1.rmin, o1.rmax] + list(rp[1][2:]) + list(rp[2][2:])
            edges2 = [o2.rmin, o2.rmax] + list(rp[1][:2]) + list(rp[2][:2])
            far = np.ones(coords.shape[:2], bool)
            for e in edges1:
                far &= np.abs(coords[..., 0] - e) > 4 * h
            for e in edges2:
                far &= np.abs(coords[..., 1] - e) > 4 * h
            assert np.abs(gx[..., k] - fd)[far].max() <= 1e-6 * max(1.0, np.abs(fd).max()), k
    assert np.array_equal(gx_a.cpu().numpy(), gx_b.cpu().numpy())
