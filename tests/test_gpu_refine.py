"""GPU parity of the mixed-precision build: FP32 sweeps with FP64 residual
refinement (builder precision FP32_REFINE) against the FP64 oracle.

Bar: on every line, max |c - c_ref| / max |c_ref| <= 1e-12 (the precision
form's stated target), including the geometrically stretched Hermite meshes
and lines scaled by powers of two far outside FP32's range.
"""
import numpy as np
import pytest

from cases import STRETCHED_CASES, TILED_CASES, Case, G, H, P, box_filtered_normal, derivs, values

TOL = 1e-12
FP32_REFINE = 1

pytestmark = pytest.mark.gpu


def _line_err(got, ref, axis):
    num = np.abs(got - ref).max(axis=axis)
    den = np.abs(ref).max(axis=axis)
    return np.max(np.where(den > 0, num / np.where(den > 0, den, 1), num))


def _run(case, nbatch, layout, cuda, seed=0, scales=None, smooth=False, steps=0):
    import torch

    orc = case.oracle()
    bs, builder = case.product(precision=FP32_REFINE, refine_steps=steps)
    assert builder.info.kernel == 2, "mixed-precision kernel not selected"
    assert builder.info.refine_steps in (1, 2)
    if steps:
        assert builder.info.refine_steps == steps
    npts = orc.npts
    f = box_filtered_normal(npts, nbatch, seed) if smooth else values(npts, nbatch, seed, x=orc.points())
    dl = derivs(orc.nbv_l, nbatch, seed) if case.bcl == H else None
    dr = derivs(orc.nbv_r, nbatch, seed + 7) if case.bcu == H else None
    if scales is not None:
        f = f * scales[None, :]
        dl = None if dl is None else dl * scales[None, :]
        dr = None if dr is None else dr * scales[None, :]
    if layout == "batch_contiguous":
        ref = orc.build(f, dl, dr)
        coef = torch.full((orc.size, nbatch), np.nan, dtype=torch.float64, device=cuda)
        tdl = torch.from_numpy(dl).to(cuda) if dl is not None else None
        tdr = torch.from_numpy(dr).to(cuda) if dr is not None else None
        builder(coef, torch.from_numpy(f).to(cuda), tdl, tdr, axis=0)
        got, ax = coef.cpu().numpy(), 0
    else:
        ref = orc.build(f, dl, dr)
        coef = torch.full((nbatch, orc.size), np.nan, dtype=torch.float64, device=cuda)
        tdl = torch.from_numpy(np.ascontiguousarray(dl.T)).to(cuda) if dl is not None else None
        tdr = torch.from_numpy(np.ascontiguousarray(dr.T)).to(cuda) if dr is not None else None
        builder(coef, torch.from_numpy(np.ascontiguousarray(f.T)).to(cuda), tdl, tdr, axis=1)
        got, ax = coef.cpu().numpy().T, 0
    assert np.all(np.isfinite(got)), "coefficients not fully written"
    return got, ref, ax, builder


@pytest.mark.parametrize("case", STRETCHED_CASES, ids=lambda c: c.id)
@pytest.mark.parametrize("layout", ["batch_contiguous", "interp_contiguous"])
def test_refine_stretched_mesh(case, layout, cuda):
    got, ref, ax, b = _run(case, 48, layout, cuda, smooth=True)
    err = _line_err(got, ref, ax)
    print(f"{case.id} {layout}: steps {b.info.refine_steps} contraction {b.info.refine_contraction:.2e} err {err:.2e}")
    assert err <= TOL


@pytest.mark.parametrize("case", TILED_CASES, ids=lambda c: c.id)
def test_refine_matches_oracle(case, cuda):
    got, ref, ax, _ = _run(case, 40, "batch_contiguous", cuda, seed=1)
    assert _line_err(got, ref, ax) <= TOL


def test_refine_line_scaling(cuda):
    """Lines scaled by 2^-1000 .. 2^700 and 1e-30 / 1e30, and an all-zero line.
    (On this mesh the Hermite border entries reach ~2^140, so the FP64 solve
    itself overflows for data near 2^900; 2^700 is the top of the range.)"""
    case = STRETCHED_CASES[0]
    nbatch = 9
    scales = np.array([2.0 ** -1000, 2.0 ** -300, 1e-30, 1.0, 1e30, 2.0 ** 300, 2.0 ** 700, 0.0, 3.0])
    got, ref, ax, _ = _run(case, nbatch, "batch_contiguous", cuda, seed=5, scales=scales, smooth=True)
    assert np.all(np.isfinite(ref))
    assert np.all(got[:, 7] == 0.0)
    assert _line_err(got, ref, ax) <= TOL


@pytest.mark.parametrize("nbatch", [1, 7, 8, 9, 300])
def test_refine_ragged_batches(nbatch, cuda):
    got, ref, ax, _ = _run(Case(3, False, G, ncells=1024, mesh="random"), nbatch, "batch_contiguous", cuda)
    assert _line_err(got, ref, ax) <= TOL


@pytest.mark.parametrize("case", [STRETCHED_CASES[0], Case(5, True, P, ncells=1024)], ids=lambda c: c.id)
def test_refine_two_steps(case, cuda):
    """Two refinement steps requested explicitly (the 2-step kernels)."""
    got, ref, ax, b = _run(case, 33, "batch_contiguous", cuda, smooth=True, steps=2)
    assert _line_err(got, ref, ax) <= TOL
