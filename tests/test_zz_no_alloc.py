"""The allocation-free entry point ddcx_evaluator2d_eval_vjp_binned_static
(workspace in a static device array of the library) against the existing
entry point ddcx_evaluator2d_eval_vjp_binned with a caller workspace: two
calls in a row, the second with fewer points and batch entries, each
compared with the workspace form on the same inputs (coordinate gradients
bit for bit, coefficient gradients to 1e-12: both accumulate with LDS
atomics in a scheduling-dependent order); sizes beyond the static
workspace are refused."""
import ctypes as C

import numpy as np
import pytest

from cases import breaks_for

pytestmark = pytest.mark.gpu


def _pair():
    import ddc_amd as D

    bs1 = D.NonUniformBSplines(3, breaks_for("random", 70, seed=5))
    bs2 = D.UniformBSplines(2, 0.0, 1.0, 45, periodic=True)
    return D.SplineEvaluator2D(
        D.SplineEvaluator(bs1, D.ConstantExtrapolationRule(0.2), D.ConstantExtrapolationRule(0.9)),
        D.SplineEvaluator(bs2, D.PeriodicExtrapolationRule(), D.PeriodicExtrapolationRule())), bs1, bs2


def _inputs(bs1, bs2, nb, npts, seed, cuda):
    import torch

    rng = np.random.default_rng(seed)
    coef = torch.from_numpy(rng.standard_normal((nb, bs1.size(), bs2.size()))).to(cuda)
    xy = torch.from_numpy(np.stack([rng.uniform(-0.2, 1.2, (nb, npts)), rng.uniform(-0.3, 1.3, (nb, npts))],
                                   axis=-1)).to(cuda)
    w = torch.from_numpy(rng.standard_normal((nb, npts))).to(cuda)
    return coef, xy, w


def _call(name, ev, xy, w, coef, gx, gc, extra=()):
    from ddc_amd._lib import call

    nb, npts = w.shape
    call(name, ev.ev1._h, ev.ev2._h, xy.data_ptr(), w.data_ptr(), gx.data_ptr(), npts, coef.data_ptr(),
         gc.data_ptr(), coef.stride(0), coef.stride(1), coef.stride(2), nb, *extra, None)


def test_static_workspace_matches_caller_workspace(cuda):
    import torch

    ev, bs1, bs2 = _pair()
    for nb, npts, seed in ((16, 20000, 1), (3, 700, 2)):
        coef, xy, w = _inputs(bs1, bs2, nb, npts, seed, cuda)
        need = C.c_int64(0)
        from ddc_amd._lib import call

        call("ddcx_evaluator2d_eval_vjp_workspace", ev.ev1._h, ev.ev2._h, npts, nb, C.byref(need))
        ws = torch.empty(need.value, dtype=torch.uint8, device=cuda)
        gx_a, gc_a = torch.full_like(xy, np.nan), torch.full_like(coef, np.nan)
        _call("ddcx_evaluator2d_eval_vjp_binned", ev, xy, w, coef, gx_a, gc_a, (ws.data_ptr(), ws.numel()))
        gx_s, gc_s = torch.full_like(xy, np.nan), torch.full_like(coef, np.nan)
        _call("ddcx_evaluator2d_eval_vjp_binned_static", ev, xy, w, coef, gx_s, gc_s)
        torch.cuda.synchronize()
        a, s = gc_a.cpu().numpy(), gc_s.cpu().numpy()
        assert np.all(np.isfinite(s)) and np.all(np.isfinite(gx_s.cpu().numpy()))
        assert np.abs(s - a).max() <= 1e-12 * np.abs(a).max()
        assert np.array_equal(gx_s.cpu().numpy(), gx_a.cpu().numpy())


def test_static_workspace_limit(cuda):
    import torch
    from ddc_amd._lib import DDCXError, lib

    ev, bs1, bs2 = _pair()
    nb, npts = 16, 1 << 21  # needs ~128 MiB of workspace
    need = C.c_int64(0)
    from ddc_amd._lib import call

    call("ddcx_evaluator2d_eval_vjp_workspace", ev.ev1._h, ev.ev2._h, npts, nb, C.byref(need))
    assert need.value > (64 << 20)
    coef = torch.zeros((nb, bs1.size(), bs2.size()), dtype=torch.float64, device=cuda)
    xy = torch.zeros((nb, 1, 2), dtype=torch.float64, device=cuda)  # never read: refused before any launch
    with pytest.raises(DDCXError, match="static workspace"):
        call("ddcx_evaluator2d_eval_vjp_binned_static", ev.ev1._h, ev.ev2._h, xy.data_ptr(), xy.data_ptr(),
             xy.data_ptr(), npts, coef.data_ptr(), coef.data_ptr(), coef.stride(0), coef.stride(1), coef.stride(2),
             nb, None)
    del lib
