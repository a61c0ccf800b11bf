"""torch.autograd wrappers of the spline build and evaluation.

``SplineBuild`` differentiates SplineBuilder::operator()
(spline_builder.hpp:538-557) with respect to the values and the Hermite
boundary derivatives; its backward is ``ddcx_builder_build_vjp`` (the
transposed block solve).  ``SplineEval`` differentiates
SplineEvaluator::operator() (spline_evaluator.hpp:257-294) with respect to
the coefficients and the evaluation coordinates; its backward is
``ddcx_evaluator_eval_vjp`` (scatter-add of basis weights, and w * f'(x)).
Both run the HIP kernels through the C ABI on torch's current stream, in
FP64; neither is differentiable twice.

``interpolate(builder, evaluator, vals, coords)`` composes them: the spline
of ``vals`` evaluated at ``coords``, differentiable in both (the
semi-Lagrangian advection step's dL/dvals and dL/dfeet).
"""
from __future__ import annotations

import torch

from .splines import SplineBuilder, SplineEvaluator


class SplineBuild(torch.autograd.Function):
    """coef = builder(vals, derivs_xmin, derivs_xmax) along ``axis``."""

    @staticmethod
    def forward(ctx, builder: SplineBuilder, vals, derivs_xmin=None, derivs_xmax=None, axis=0):
        shape = list(vals.shape)
        shape[axis] = builder.info.size
        coef = torch.empty(shape, dtype=vals.dtype, device=vals.device)
        builder(coef, vals, derivs_xmin, derivs_xmax, axis=axis)
        ctx.builder, ctx.axis = builder, axis
        ctx.vals_shape = vals.shape
        ctx.dshape = [None if d is None else d.shape for d in (derivs_xmin, derivs_xmax)]
        return coef

    @staticmethod
    def backward(ctx, gcoef):
        b = ctx.builder
        gcoef = gcoef.contiguous()
        gvals = torch.empty(ctx.vals_shape, dtype=gcoef.dtype, device=gcoef.device)
        gd = [None if s is None else torch.empty(s, dtype=gcoef.dtype, device=gcoef.device) for s in ctx.dshape]
        b.build_vjp(gcoef, gvals, gd[0], gd[1], axis=ctx.axis)
        return None, gvals, gd[0], gd[1], None


class SplineEval(torch.autograd.Function):
    """out = evaluator(coords, coef) along ``axis`` (coords and out share a
    shape; coef has size() rows along ``axis``)."""

    @staticmethod
    def forward(ctx, evaluator: SplineEvaluator, coords, coef, axis=0):
        coords, coef = coords.contiguous(), coef.contiguous()
        out = torch.empty_like(coords)
        evaluator(out, coords, coef, axis=axis)
        ctx.evaluator, ctx.axis = evaluator, axis
        ctx.save_for_backward(coords, coef)
        return out

    @staticmethod
    def backward(ctx, gout):
        coords, coef = ctx.saved_tensors
        w = gout.contiguous()
        gx = torch.empty_like(coords)
        gcoef = torch.empty_like(coef)
        ctx.evaluator.eval_vjp(coords, w, coef, gx, gcoef, axis=ctx.axis)
        return None, gx, gcoef, None


def interpolate(builder: SplineBuilder, evaluator: SplineEvaluator, vals, coords, derivs_xmin=None,
                derivs_xmax=None, axis=0):
    """The spline of ``vals`` (interpolation points along ``axis``) evaluated
    at ``coords`` (same batch dimensions), differentiable in vals, the
    Hermite derivatives and coords."""
    if evaluator.bsplines is not builder.bsplines:
        raise ValueError("the evaluator must use the builder's basis")
    coef = SplineBuild.apply(builder, vals, derivs_xmin, derivs_xmax, axis)
    return SplineEval.apply(evaluator, coords, coef, axis)
