"""GPU parity of the adjoint entry points against the oracle's adjoints.

build_vjp: <= 1e-13 relative (max norm).  eval_vjp: coordinate gradients
<= 1e-13; coefficient gradients <= 1e-12 (the GPU accumulates them with
atomics, in an order that differs from the oracle's sequential sum).
"""
import numpy as np
import pytest

from cases import ALL_CASES, PIVOTED_CASES, SHORT_TILED_CASES, TILED_CASES, Case, G, H, P
from conftest import rel_err

pytestmark = pytest.mark.gpu


def _vjp(case, nbatch, layout, cuda, seed=0):
    import torch

    orc = case.oracle()
    bs, builder = case.product()
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((orc.size, nbatch))
    gv_ref, gdl_ref, gdr_ref = orc.build_vjp(g)
    nl, nr = orc.nbv_l, orc.nbv_r
    if layout == "batch_contiguous":
        tg = torch.from_numpy(g).to(cuda)
        gv = torch.full((orc.npts, nbatch), np.nan, dtype=torch.float64, device=cuda)
        gdl = torch.full((nl, nbatch), np.nan, dtype=torch.float64, device=cuda) if nl else None
        gdr = torch.full((nr, nbatch), np.nan, dtype=torch.float64, device=cuda) if nr else None
        builder.build_vjp(tg, gv, gdl, gdr, axis=0)
        got = [gv.cpu().numpy()] + ([gdl.cpu().numpy()] if nl else []) + ([gdr.cpu().numpy()] if nr else [])
    else:
        tg = torch.from_numpy(np.ascontiguousarray(g.T)).to(cuda)
        gv = torch.full((nbatch, orc.npts), np.nan, dtype=torch.float64, device=cuda)
        gdl = torch.full((nbatch, nl), np.nan, dtype=torch.float64, device=cuda) if nl else None
        gdr = torch.full((nbatch, nr), np.nan, dtype=torch.float64, device=cuda) if nr else None
        builder.build_vjp(tg, gv, gdl, gdr, axis=1)
        got = [gv.cpu().numpy().T] + ([gdl.cpu().numpy().T] if nl else []) + ([gdr.cpu().numpy().T] if nr else [])
    ref = [gv_ref] + ([gdl_ref] if nl else []) + ([gdr_ref] if nr else [])
    return np.concatenate(got), np.concatenate(ref)


@pytest.mark.parametrize("case", ALL_CASES, ids=lambda c: c.id)
@pytest.mark.parametrize("layout", ["batch_contiguous", "interp_contiguous"])
def test_build_vjp_matches_oracle(case, layout, cuda):
    got, ref = _vjp(case, 13, layout, cuda)
    assert np.all(np.isfinite(got))
    assert rel_err(got, ref) <= 1e-13


@pytest.mark.parametrize("case", PIVOTED_CASES, ids=lambda c: c.id)
@pytest.mark.parametrize("layout", ["batch_contiguous", "interp_contiguous"])
def test_build_vjp_pivoted_band(case, layout, cuda):
    """The transposed solve with the interchanges applied in reverse (dgbtrs trans='T')."""
    _, builder = case.product()
    assert builder.info.pivoted == 1
    got, ref = _vjp(case, 29, layout, cuda, seed=3)
    assert np.all(np.isfinite(got))
    assert rel_err(got, ref) <= 1e-13


@pytest.mark.parametrize("case", TILED_CASES + SHORT_TILED_CASES, ids=lambda c: c.id)
@pytest.mark.parametrize("layout", ["batch_contiguous", "interp_contiguous"])
def test_build_vjp_tiled(case, layout, cuda):
    """Main blocks of 64..2048 rows: the tiled transposed solve."""
    got, ref = _vjp(case, 40, layout, cuda, seed=2)
    assert np.all(np.isfinite(got))
    assert rel_err(got, ref) <= 1e-13


@pytest.mark.parametrize("case", [c for c in ALL_CASES if c.degree in (2, 3, 5)], ids=lambda c: c.id)
def test_eval_vjp_matches_oracle(case, cuda):
    import torch
    import ddc_amd as D

    orc = case.oracle()
    bs, _ = case.product()
    rng = np.random.default_rng(7)
    nbatch, npts = 6, 80
    coef = rng.standard_normal((orc.size, nbatch))
    if case.periodic:
        coef[orc.nbasis:] = coef[: case.degree]
    L = orc.rmax - orc.rmin
    x = rng.uniform(orc.rmin - 0.1 * L, orc.rmax + 0.1 * L, (npts, nbatch))
    w = rng.standard_normal((npts, nbatch))
    gc_ref, gx_ref = orc.eval_vjp(x, w, coef, extrap=("constant", "null"))
    ev = D.SplineEvaluator(bs, D.ConstantExtrapolationRule(), D.NullExtrapolationRule())
    tx, tw, tc = (torch.from_numpy(a).to(cuda) for a in (x, w, coef))
    gx = torch.empty_like(tx)
    gc = torch.empty_like(tc)
    ev.eval_vjp(tx, tw, tc, gx, gc)
    assert rel_err(gx.cpu().numpy(), gx_ref) <= 1e-13
    assert rel_err(gc.cpu().numpy(), gc_ref) <= 1e-12


@pytest.mark.parametrize("case", [Case(3, True, P, ncells=1024), Case(5, True, P, ncells=300),
                                  Case(3, False, G, ncells=1024, mesh="random"),
                                  Case(4, False, H, ncells=500, mesh="random"), Case(2, True, G, ncells=200)],
                         ids=lambda c: c.id)
@pytest.mark.parametrize("layout", ["batch_contiguous", "interp_contiguous"])
def test_eval_vjp_tiled(case, layout, cuda):
    """LDS-accumulated evaluator adjoint (>= 8 lines, >= 64 points), with
    points outside the domain under the constant and null rules."""
    import torch
    import ddc_amd as D

    orc = case.oracle()
    bs, _ = case.product()
    rng = np.random.default_rng(9)
    nbatch, npts = 21, 300
    coef = rng.standard_normal((orc.size, nbatch))
    if case.periodic:
        coef[orc.nbasis:] = coef[: case.degree]
    L = orc.rmax - orc.rmin
    x = rng.uniform(orc.rmin - 0.1 * L, orc.rmax + 0.1 * L, (npts, nbatch))
    w = rng.standard_normal((npts, nbatch))
    gc_ref, gx_ref = orc.eval_vjp(x, w, coef, extrap=("constant", "null"))
    ev = D.SplineEvaluator(bs, D.ConstantExtrapolationRule(), D.NullExtrapolationRule())
    if layout == "batch_contiguous":
        tx, tw, tc = (torch.from_numpy(a).to(cuda) for a in (x, w, coef))
        gx, gc = torch.full_like(tx, np.nan), torch.full_like(tc, np.nan)
        ev.eval_vjp(tx, tw, tc, gx, gc, axis=0)
        gx, gc = gx.cpu().numpy(), gc.cpu().numpy()
    else:
        tx, tw, tc = (torch.from_numpy(np.ascontiguousarray(a.T)).to(cuda) for a in (x, w, coef))
        gx, gc = torch.full_like(tx, np.nan), torch.full_like(tc, np.nan)
        ev.eval_vjp(tx, tw, tc, gx, gc, axis=1)
        gx, gc = gx.cpu().numpy().T, gc.cpu().numpy().T
    assert np.all(np.isfinite(gc)) and np.all(np.isfinite(gx))
    assert rel_err(gx, gx_ref) <= 1e-13
    assert rel_err(gc, gc_ref) <= 1e-12
