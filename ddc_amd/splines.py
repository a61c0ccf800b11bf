"""Host-side mirror of DDC's spline interface over the libddcx C ABI.

Class and method names follow the reference (src/ddc/kernels/splines/):
``UniformBSplines`` / ``NonUniformBSplines`` (bsplines_uniform.hpp:48-370,
bsplines_non_uniform.hpp:44-392), ``SplineBuilder`` (spline_builder.hpp:57-617),
``SplineEvaluator`` (spline_evaluator.hpp:33-606), ``SplineBuilder2D`` /
``SplineEvaluator2D`` (spline_builder_2d.hpp:27-651,
spline_evaluator_2d.hpp:37-826), ``SplineBuilderClosure``, ``SplineSolver``,
the extrapolation rules, ``GrevilleInterpolationPoints`` and ``integrals``.

Arrays are torch tensors in device memory.  The reference names the
interpolation dimension by type; here it is named by ``axis`` (default 0).
Any other dimensions (at most two) are batch dimensions, in any order.  All
calls are enqueued on torch's current stream.
"""
from __future__ import annotations

import ctypes as C
import enum

import numpy as np
import torch

from ._lib import (call, ddcx_builder_desc, ddcx_builder_info, ddcx_bsplines_desc, ddcx_extrap_rule, ddcx_layout,
                   lib)


class SplineBuilderClosure(enum.IntEnum):
    """ddc::SplineBuilderClosure, spline_builder_closures.hpp:18-23."""

    PERIODIC = 0
    HERMITE = 1
    HOMOGENEOUS_HERMITE = 2
    GREVILLE = 3


class SplineSolver(enum.IntEnum):
    """ddc::SplineSolver, spline_builder.hpp:36-39 (GINKGO routes to the direct solver)."""

    GINKGO = 0
    LAPACK = 1


class Precision(enum.IntEnum):
    FP64 = 0
    FP32_REFINE = 1


_NAN = float("nan")


class NullExtrapolationRule:
    """null_extrapolation_rule.hpp:14-26: 0 outside the domain."""

    code = 0

    def _desc(self):
        return ddcx_extrap_rule(self.code, _NAN, _NAN, _NAN)


class ConstantExtrapolationRule:
    """ddc::ConstantExtrapolationRule (constant_extrapolation_rule.hpp:27-195):
    outside the domain the spline takes its value at ``eval_pos``, a
    coordinate inside the domain (ConstantExtrapolationRule<DimI>(eval_pos),
    :39-40).  In a SplineEvaluator2D, the two-dimensional form
    ConstantExtrapolationRule<DimI, DimNI>(eval_pos, min, max) (:99-124) also
    clamps the other coordinate to [eval_pos_not_interest_min,
    eval_pos_not_interest_max] when the other dimension is not periodic.
    ``eval_pos=None`` evaluates at the bound the rule guards (rmin for a lower
    rule, rmax for an upper one); a None pair clamps to the other dimension's
    domain."""

    code = 1

    def __init__(self, eval_pos=None, eval_pos_not_interest_min=None, eval_pos_not_interest_max=None):
        if (eval_pos_not_interest_min is None) != (eval_pos_not_interest_max is None):
            raise ValueError("give both eval_pos_not_interest_min and eval_pos_not_interest_max, or neither")
        self.eval_pos = eval_pos
        self.eval_pos_not_interest_min = eval_pos_not_interest_min
        self.eval_pos_not_interest_max = eval_pos_not_interest_max

    def _desc(self):
        def f(v):
            return _NAN if v is None else float(v)

        return ddcx_extrap_rule(self.code, f(self.eval_pos), f(self.eval_pos_not_interest_min),
                                f(self.eval_pos_not_interest_max))


class PeriodicExtrapolationRule:
    """periodic_extrapolation_rule.hpp:12-24 (the periodic wrap applies instead)."""

    code = 2

    def _desc(self):
        return ddcx_extrap_rule(self.code, _NAN, _NAN, _NAN)


def n_boundary_equations(bc, degree):
    """spline_builder_closures.hpp:33-55."""
    return degree // 2 if bc in (SplineBuilderClosure.HERMITE, SplineBuilderClosure.HOMOGENEOUS_HERMITE) else 0


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _check_f64_cuda(t, name):
    if t.dtype != torch.float64:
        raise TypeError(f"{name} must be float64")
    if not t.is_cuda:
        raise TypeError(f"{name} must be a device tensor")


def batched_layout(t, axis):
    """ddcx_layout of a tensor whose dimension `axis` is the row dimension and
    whose other dimensions are batch dimensions, enumerated in row-major order.
    Neighbouring batch dimensions on the same side of `axis` that are
    contiguous with each other (stride[d] == shape[d+1] * stride[d+1]) merge
    into one; at most two batch dimensions may remain.  Returns
    (layout, nbatch)."""
    axis = axis % t.dim()
    groups = []  # [count, stride] of merged batch dimensions, outermost first
    for side in (range(axis), range(axis + 1, t.dim())):
        cur = None
        for d in side:
            n, st = t.shape[d], t.stride(d)
            if cur is not None and cur[1] == n * st:
                cur = [cur[0] * n, st]
            else:
                if cur is not None:
                    groups.append(cur)
                cur = [n, st]
        if cur is not None:
            groups.append(cur)
    if len(groups) > 2:
        raise ValueError("at most two batch dimensions are supported (after merging contiguous ones)")
    rs = t.stride(axis)
    if len(groups) == 0:
        return ddcx_layout(rs, 1, 0, 0), 1
    if len(groups) == 1:
        n, st = groups[0]
        return ddcx_layout(rs, n, st, 0), n
    (n0, s0), (n1, s1) = groups
    return ddcx_layout(rs, n1, s1, s0), n0 * n1


class _BSplinesBase:
    def __init__(self, desc, keep=None):
        h = C.c_void_p()
        call("ddcx_bsplines_create", C.byref(desc), C.byref(h))
        self._h = h
        self._keep = keep
        self._degree = desc.degree
        self._ncells = desc.ncells
        self._periodic = bool(desc.periodic)
        k = np.empty(self._ncells + 1 + 2 * self._degree)
        call("ddcx_bsplines_knots", self._h, C.c_void_p(k.ctypes.data))
        self._knots = k

    def __del__(self):
        if getattr(self, "_h", None) and lib is not None:
            try:
                lib().ddcx_bsplines_destroy(self._h)
            except Exception:
                pass
            self._h = None

    def degree(self):
        return self._degree

    def is_periodic(self):
        return self._periodic

    def ncells(self):
        return self._ncells

    def nbasis(self):
        return self._ncells + (0 if self._periodic else self._degree)

    def size(self):
        return self._ncells + self._degree

    def knots(self):
        return self._knots.copy()

    def rmin(self):
        return float(self._knots[self._degree])

    def rmax(self):
        return float(self._knots[self._degree + self._ncells])

    def length(self):
        return self.rmax() - self.rmin()


class UniformBSplines(_BSplinesBase):
    """ddc::UniformBSplines<CDim, D> on [rmin, rmax] with ncells cells."""

    def __init__(self, degree, rmin, rmax, ncells, periodic=False):
        super().__init__(ddcx_bsplines_desc(degree, 1, int(periodic), ncells, float(rmin), float(rmax), None))

    def is_uniform(self):
        return True


class NonUniformBSplines(_BSplinesBase):
    """ddc::NonUniformBSplines<CDim, D> on the given sorted break points."""

    def __init__(self, degree, breaks, periodic=False):
        br = np.ascontiguousarray(np.asarray(breaks, dtype=np.float64))
        super().__init__(ddcx_bsplines_desc(degree, 0, int(periodic), br.size - 1, 0.0, 0.0,
                                            C.c_void_p(br.ctypes.data)), keep=br)

    def is_uniform(self):
        return False


def knots_as_interpolation_points(bsplines, lower, upper):
    """ddc::KnotsAsInterpolationPoints<BSplines, SBCLower, SBCUpper>
    (knots_as_interpolation_points.hpp:36-92): the break points as
    interpolation points, ncells + 1 of them (ncells for a periodic basis,
    get_domain :99-106).  Uniform bases sample them as
    UniformPointSampling::init(rmin, rmax, ncells + 1) does
    (uniform_point_sampling.hpp:141-149): rmin + i * ((rmax - rmin) / ncells).
    Not for Greville closures.  Pass the result as SplineBuilder(points=...)."""
    if SplineBuilderClosure.GREVILLE in (SplineBuilderClosure(lower), SplineBuilderClosure(upper)):
        raise ValueError("KnotsAsInterpolationPoints does not apply to Greville closures")
    n = bsplines.ncells() + (0 if bsplines.is_periodic() else 1)
    if bsplines.is_uniform():
        step = (bsplines.rmax() - bsplines.rmin()) / bsplines.ncells()
        return bsplines.rmin() + np.arange(n) * step
    d = bsplines.degree()
    return bsplines.knots()[d: d + n].copy()


def integrals(bsplines):
    """ddc::integrals(), integrals.hpp:165-177 (host array of size() values)."""
    out = np.empty(bsplines.size())
    call("ddcx_bsplines_integrals", bsplines._h, C.c_void_p(out.ctypes.data))
    return out


class SplineBuilder:
    """ddc::SplineBuilder<Exec, Mem, BSplines, DDimI, BCl, BCu, Solver>."""

    def __init__(self, bsplines, lower=None, upper=None, solver=SplineSolver.LAPACK, points=None,
                 precision=Precision.FP64, refine_steps=0):
        if lower is None:
            lower = SplineBuilderClosure.PERIODIC if bsplines.is_periodic() else SplineBuilderClosure.GREVILLE
        if upper is None:
            upper = lower
        self.bsplines = bsplines
        self.lower, self.upper = SplineBuilderClosure(lower), SplineBuilderClosure(upper)
        pts = None
        if points is not None:
            pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64))
        desc = ddcx_builder_desc(int(self.lower), int(self.upper), int(solver), int(precision),
                                 C.c_void_p(pts.ctypes.data) if pts is not None else None,
                                 0 if pts is None else pts.size, int(refine_steps))
        h = C.c_void_p()
        call("ddcx_builder_create", bsplines._h, C.byref(desc), C.byref(h))
        self._h = h
        info = ddcx_builder_info()
        call("ddcx_builder_get_info", self._h, C.byref(info))
        self.info = info
        self.s_odd = bsplines.degree() % 2
        self.s_nbe_xmin = n_boundary_equations(self.lower, bsplines.degree())
        self.s_nbe_xmax = n_boundary_equations(self.upper, bsplines.degree())
        self.s_nbv_xmin = info.nbv_lower
        self.s_nbv_xmax = info.nbv_upper

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                lib().ddcx_builder_destroy(self._h)
            except Exception:
                pass
            self._h = None

    def interpolation_points(self):
        p = np.empty(self.info.npoints)
        call("ddcx_builder_points", self._h, C.c_void_p(p.ctypes.data))
        return p

    def npoints(self):
        return self.info.npoints

    def offset(self):
        return self.info.offset

    def _layouts(self, spline, vals, derivs_xmin, derivs_xmax, axis, names=("spline", "vals", "derivs_xmin",
                                                                             "derivs_xmax")):
        """Checks the arrays of one build (or its adjoint) and returns
        (coef layout, vals layout, derivs layout, nbatch).  Every check runs
        before any launch: the kernels index every row of every line.
        spline may be None (coefficients kept on chip by the fused kernels)."""
        _check_f64_cuda(vals, names[1])
        if vals.shape[axis] != self.info.npoints:
            raise ValueError(f"{names[1]} has {vals.shape[axis]} interpolation points, expected {self.info.npoints}")
        vl, nbatch = batched_layout(vals, axis)
        cl = ddcx_layout(1, 1, 0, 0)
        if spline is not None:
            _check_f64_cuda(spline, names[0])
            if vals.dim() != spline.dim():
                raise ValueError(f"{names[0]} and {names[1]} must have the same number of dimensions")
            if spline.shape[axis] != self.info.size:
                raise ValueError(f"{names[0]} has {spline.shape[axis]} coefficients, expected {self.info.size}")
            cl, nb2 = batched_layout(spline, axis)
            if nb2 != nbatch:
                raise ValueError(f"{names[0]} and {names[1]} batch domains differ")
        dl = ddcx_layout(1, 1, 0, 0)
        layouts = []
        for d, name, nbv in ((derivs_xmin, names[2], self.s_nbv_xmin), (derivs_xmax, names[3], self.s_nbv_xmax)):
            if d is None:
                if nbv > 0 and (self.lower if name == names[2] else self.upper) == SplineBuilderClosure.HERMITE:
                    raise ValueError(f"{name} is required by the Hermite closure")
                continue
            _check_f64_cuda(d, name)
            if d.dim() != vals.dim() or d.shape[axis] < nbv:
                raise ValueError(f"{name} must have the dimensions of {names[1]} with {nbv} rows along axis")
            dl, nb3 = batched_layout(d, axis)
            if nb3 != nbatch:
                raise ValueError(f"{name} batch domain differs")
            layouts.append((dl.row_stride, dl.inner_count, dl.inner_stride, dl.outer_stride))
        if len(layouts) == 2 and layouts[0] != layouts[1]:
            raise ValueError(f"{names[2]} and {names[3]} must have the same layout")
        return cl, vl, dl, nbatch

    def __call__(self, spline, vals, derivs_xmin=None, derivs_xmax=None, axis=0):
        """SplineBuilder::operator(), spline_builder.hpp:538-557."""
        cl, vl, dl, nbatch = self._layouts(spline, vals, derivs_xmin, derivs_xmax, axis)
        call("ddcx_builder_build", self._h, _ptr(spline), cl, _ptr(vals), vl, _ptr(derivs_xmin), _ptr(derivs_xmax),
             dl, nbatch, _stream())
        return spline

    def build_vjp(self, gcoef, gvals, gderivs_xmin=None, gderivs_xmax=None, axis=0):
        """Adjoint of __call__: gvals (and Hermite derivative gradients) from gcoef."""
        cl, vl, dl, nbatch = self._layouts(gcoef, gvals, gderivs_xmin, gderivs_xmax, axis,
                                           ("gcoef", "gvals", "gderivs_xmin", "gderivs_xmax"))
        call("ddcx_builder_build_vjp", self._h, _ptr(gcoef), cl, _ptr(gvals), vl, _ptr(gderivs_xmin),
             _ptr(gderivs_xmax), dl, nbatch, _stream())
        return gvals

    def quadrature_coefficients(self):
        """SplineBuilder::quadrature_coefficients(), spline_builder.hpp:1111-1234 (host arrays)."""
        qa = np.empty(max(self.s_nbv_xmin, 1))
        qv = np.empty(self.info.nbasis - self.s_nbv_xmin - self.s_nbv_xmax)
        qb = np.empty(max(self.s_nbv_xmax, 1))
        call("ddcx_builder_quadrature", self._h, C.c_void_p(qa.ctypes.data), C.c_void_p(qv.ctypes.data),
             C.c_void_p(qb.ctypes.data))
        return qa[: self.s_nbv_xmin], qv, qb[: self.s_nbv_xmax]


class SplineEvaluator:
    """ddc::SplineEvaluator<Exec, Mem, BSplines, EvalDDim, LowerRule, UpperRule>."""

    def __init__(self, bsplines, lower_extrap_rule=None, upper_extrap_rule=None):
        lo = lower_extrap_rule if lower_extrap_rule is not None else (
            PeriodicExtrapolationRule() if bsplines.is_periodic() else NullExtrapolationRule())
        up = upper_extrap_rule if upper_extrap_rule is not None else lo
        self.bsplines = bsplines
        self.lower_extrap_rule, self.upper_extrap_rule = lo, up
        h = C.c_void_p()
        dlo, dup = lo._desc(), up._desc()
        call("ddcx_evaluator_create_with_rules", bsplines._h, C.byref(dlo), C.byref(dup), C.byref(h))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                lib().ddcx_evaluator_destroy(self._h)
            except Exception:
                pass
            self._h = None

    def _args(self, spline_eval, coords_eval, spline_coef, axis):
        _check_f64_cuda(spline_eval, "spline_eval")
        _check_f64_cuda(coords_eval, "coords_eval")
        _check_f64_cuda(spline_coef, "spline_coef")
        if spline_eval.shape != coords_eval.shape or spline_eval.stride() != coords_eval.stride():
            raise ValueError("spline_eval and coords_eval must have the same shape and layout")
        pl, nbatch = batched_layout(coords_eval, axis)
        cl, nb2 = batched_layout(spline_coef, axis)
        if nb2 != nbatch:
            raise ValueError("coordinates and coefficients batch domains differ")
        if spline_coef.shape[axis] != self.bsplines.size():
            raise ValueError("spline_coef has the wrong number of coefficients")
        return pl, coords_eval.shape[axis], cl, nbatch

    def __call__(self, spline_eval, coords_eval, spline_coef, axis=0):
        """SplineEvaluator::operator(), spline_evaluator.hpp:257-294."""
        pl, npts, cl, nbatch = self._args(spline_eval, coords_eval, spline_coef, axis)
        call("ddcx_evaluator_eval", self._h, _ptr(spline_eval), _ptr(coords_eval), pl, npts, _ptr(spline_coef), cl,
             nbatch, _stream())
        return spline_eval

    def deriv(self, order, spline_eval, coords_eval, spline_coef, axis=0):
        """SplineEvaluator::deriv(), spline_evaluator.hpp:357-430."""
        pl, npts, cl, nbatch = self._args(spline_eval, coords_eval, spline_coef, axis)
        call("ddcx_evaluator_deriv", self._h, int(order), _ptr(spline_eval), _ptr(coords_eval), pl, npts,
             _ptr(spline_coef), cl, nbatch, _stream())
        return spline_eval

    def eval_at_points(self, spline_eval, points, spline_coef, axis=0, order=0):
        """SplineEvaluator::operator()(spline_eval, spline_coef) and
        deriv(order, spline_eval, spline_coef) (spline_evaluator.hpp:296-342,
        :432-478): the evaluation points are the evaluation domain's own
        coordinates, `points` (one per entry of spline_eval along `axis`), the
        same for every line."""
        _check_f64_cuda(spline_eval, "spline_eval")
        _check_f64_cuda(points, "points")
        _check_f64_cuda(spline_coef, "spline_coef")
        pts = points.contiguous()
        if pts.dim() != 1 or pts.numel() != spline_eval.shape[axis]:
            raise ValueError("points must hold one coordinate per entry of spline_eval along axis")
        pl, nbatch = batched_layout(spline_eval, axis)
        cl, nb2 = batched_layout(spline_coef, axis)
        if nb2 != nbatch:
            raise ValueError("evaluation and coefficient batch domains differ")
        if spline_coef.shape[axis] != self.bsplines.size():
            raise ValueError("spline_coef has the wrong number of coefficients")
        call("ddcx_evaluator_eval_at", self._h, int(order), _ptr(spline_eval), pl, _ptr(pts), pts.numel(),
             _ptr(spline_coef), cl, nbatch, _stream())
        return spline_eval

    def integrate(self, integrals_out, spline_coef, axis=0):
        """SplineEvaluator::integrate(), spline_evaluator.hpp:493-526."""
        _check_f64_cuda(integrals_out, "integrals")
        cl, nbatch = batched_layout(spline_coef, axis)
        if integrals_out.numel() != nbatch or not integrals_out.is_contiguous():
            raise ValueError("integrals must be a contiguous tensor over the batch domain")
        call("ddcx_evaluator_integrate", self._h, _ptr(integrals_out), 1, _ptr(spline_coef), cl, nbatch, _stream())
        return integrals_out

    def eval_vjp(self, coords_eval, w, spline_coef, gx, gcoef, axis=0):
        """Adjoint of __call__: gcoef = dL/dcoef, gx = dL/dcoords for L = sum(w * f(coords)).
        w and gx share the layout of coords_eval, gcoef that of spline_coef."""
        pl, npts, cl, nbatch = self._args(gx, coords_eval, spline_coef, axis)
        _check_f64_cuda(w, "w")
        _check_f64_cuda(gcoef, "gcoef")
        if w.shape != coords_eval.shape or w.stride() != coords_eval.stride():
            raise ValueError("w must have the shape and layout of coords_eval")
        if gcoef.shape != spline_coef.shape or gcoef.stride() != spline_coef.stride():
            raise ValueError("gcoef must have the shape and layout of spline_coef")
        call("ddcx_evaluator_eval_vjp", self._h, _ptr(coords_eval), _ptr(w), _ptr(gx), pl, npts,
             _ptr(spline_coef), _ptr(gcoef), cl, nbatch, _stream())
        return gcoef, gx


def build_eval(builder, evaluator, spline_eval, coords_eval, vals, spline=None, derivs_xmin=None, derivs_xmax=None,
               axis=0):
    """builder(spline, vals, ...) then evaluator(spline_eval, coords_eval, spline)
    in one pass (ddcx_build_eval).  With spline=None the coefficients stay on
    chip; that needs the fused kernel (FP64 builder, degree 2..5, batch-
    contiguous arrays)."""
    if evaluator.bsplines is not builder.bsplines:
        raise ValueError("the evaluator must use the builder's basis")
    cl, vl, dl, nbatch = builder._layouts(spline, vals, derivs_xmin, derivs_xmax, axis)
    _check_f64_cuda(spline_eval, "spline_eval")
    _check_f64_cuda(coords_eval, "coords_eval")
    if spline_eval.shape != coords_eval.shape or spline_eval.stride() != coords_eval.stride():
        raise ValueError("spline_eval and coords_eval must have the same shape and layout")
    pl, nb2 = batched_layout(coords_eval, axis)
    if nb2 != nbatch:
        raise ValueError("coords_eval and vals batch domains differ")
    call("ddcx_build_eval", builder._h, evaluator._h, _ptr(spline), cl, _ptr(vals), vl, _ptr(derivs_xmin),
         _ptr(derivs_xmax), dl, _ptr(spline_eval), _ptr(coords_eval), pl, coords_eval.shape[axis], nbatch, _stream())
    return spline_eval


def advect(builder, evaluator, out, vals, shift, spline=None, derivs_xmin=None, derivs_xmax=None, axis=0):
    """One semi-Lagrangian advection step (examples/characteristics_advection.cpp:
    237-274) in one fused kernel (ddcx_advect): out[i, l] is the spline of
    vals along `axis` evaluated at the foot x_i - shift[l] of interpolation
    point x_i, for each batch line l (shift has one value per line, in the
    batch order of vals).  The feet are computed in the kernel and the
    coefficients stay on chip unless `spline` is given.  Raises where the
    fused kernel does not apply (FP64 builder, degree 2..5, batch-contiguous
    arrays, main blocks of 256 to 2048 rows)."""
    if evaluator.bsplines is not builder.bsplines:
        raise ValueError("the evaluator must use the builder's basis")
    _check_f64_cuda(shift, "shift")
    _check_f64_cuda(out, "out")
    cl, vl, dl, nbatch = builder._layouts(spline, vals, derivs_xmin, derivs_xmax, axis)
    if out.shape != vals.shape:
        raise ValueError("out must have the shape of vals")
    ol, _ = batched_layout(out, axis)
    if shift.numel() != nbatch or not shift.is_contiguous():
        raise ValueError("shift must be a contiguous tensor with one value per batch line")
    call("ddcx_advect", builder._h, evaluator._h, _ptr(spline), cl, _ptr(vals), vl, _ptr(derivs_xmin),
         _ptr(derivs_xmax), dl, _ptr(shift), _ptr(out), ol, nbatch, _stream())
    return out


class SplineBuilder2D:
    """ddc::SplineBuilder2D (spline_builder_2d.hpp:27-651): builds along dim 1
    batched over (batch, dim 2) into a temporary, then along dim 2 batched over
    (batch, bsplines 1).  Arrays are (batch, n1, n2) (or (n1, n2))."""

    def __init__(self, builder1: SplineBuilder, builder2: SplineBuilder):
        self.builder1, self.builder2 = builder1, builder2
        self._tmp = {}

    def _scratch(self, key, shape, device):
        t = self._tmp.get(key)
        if t is None or tuple(t.shape) != tuple(shape) or t.device != device:
            t = self._tmp[key] = torch.empty(shape, dtype=torch.float64, device=device)
        return t

    def __call__(self, spline, vals, derivs_min1=None, derivs_max1=None, derivs_min2=None, derivs_max2=None,
                 mixed_derivs_min1_min2=None, mixed_derivs_max1_min2=None, mixed_derivs_min1_max2=None,
                 mixed_derivs_max1_max2=None):
        """SplineBuilder2D::operator() (spline_builder_2d.hpp:514-651).  With
        Hermite closures: derivs_*1 (batch, nbe1, n2) are the dimension-1
        derivatives at each dimension-2 point, derivs_*2 (batch, n1, nbe2) the
        dimension-2 derivatives at each dimension-1 point, mixed_* (batch,
        nbe1, nbe2) the cross derivatives at the corners.  The dimension-2
        derivatives are built along dimension 1 first (the corner derivatives
        are their Hermite data), then dimension 2 is built over the
        dimension-1 spline with them as its Hermite data."""
        squeeze = vals.dim() == 2

        def b(t):
            return None if t is None else (t.unsqueeze(0) if squeeze else t)

        v, s = b(vals), b(spline)
        nb, n1, n2 = v.shape
        b1, b2 = self.builder1, self.builder2
        s1 = b1.bsplines.size()
        H = SplineBuilderClosure.HERMITE
        dev = v.device
        d1min = d1max = None
        if b2.lower == H:
            if derivs_min2 is None:
                raise ValueError("derivs_min2 is required by the Hermite closure of dimension 2")
            dm = b(derivs_min2)
            d1min = self._scratch("dmin", (nb, s1, dm.shape[2]), dev)
            b1(d1min, dm, b(mixed_derivs_min1_min2), b(mixed_derivs_max1_min2), axis=1)
        spline1 = self._scratch("s1", (nb, s1, n2), dev)
        b1(spline1, v, b(derivs_min1), b(derivs_max1), axis=1)
        if b2.upper == H:
            if derivs_max2 is None:
                raise ValueError("derivs_max2 is required by the Hermite closure of dimension 2")
            dM = b(derivs_max2)
            d1max = self._scratch("dmax", (nb, s1, dM.shape[2]), dev)
            b1(d1max, dM, b(mixed_derivs_min1_max2), b(mixed_derivs_max1_max2), axis=1)
        b2(s, spline1, d1min, d1max, axis=2)
        return spline

    def build_vjp(self, gspline, gvals, gderivs_min1=None, gderivs_max1=None, gderivs_min2=None, gderivs_max2=None,
                  gmixed_derivs_min1_min2=None, gmixed_derivs_max1_min2=None, gmixed_derivs_min1_max2=None,
                  gmixed_derivs_max1_max2=None):
        """Adjoint of __call__: gvals (batch, n1, n2) and, with Hermite
        closures, the gradients of the derivative inputs (each shaped as the
        __call__ argument it belongs to, overwritten; None: not returned)
        from gspline (batch, size1, size2).  __call__'s passes in reverse:
        the transposed dimension-2 solve gives the gradients of the
        dimension-1 spline and of the dimension-1 splines of derivs_*2, then
        transposed dimension-1 solves give the gradients of vals, derivs_*1,
        derivs_*2 and the corner derivatives."""
        squeeze = gspline.dim() == 2

        def b(t):
            return None if t is None else (t.unsqueeze(0) if squeeze else t)

        g, gv = b(gspline), b(gvals)
        _check_f64_cuda(gv, "gvals")
        nb, n1, n2 = gv.shape
        b1, b2 = self.builder1, self.builder2
        s1 = b1.bsplines.size()
        H = SplineBuilderClosure.HERMITE
        dev = g.device

        def out(t, key, shape):
            # the caller's gradient, or scratch when it is not asked for
            return b(t) if t is not None else self._scratch(("g",) + key, shape, dev)

        lo1, hi1 = b1.lower == H, b1.upper == H
        lo2, hi2 = b2.lower == H, b2.upper == H
        nv1l, nv1u, nv2l, nv2u = b1.s_nbv_xmin, b1.s_nbv_xmax, b2.s_nbv_xmin, b2.s_nbv_xmax
        gs1 = self._scratch(("g", "s1"), (nb, s1, n2), dev)
        gd1min = self._scratch(("g", "dmin"), (nb, s1, nv2l), dev) if lo2 else None
        gd1max = self._scratch(("g", "dmax"), (nb, s1, nv2u), dev) if hi2 else None
        b2.build_vjp(g, gs1, gd1min, gd1max, axis=2)
        b1.build_vjp(gs1, gv, out(gderivs_min1, ("d1l",), (nb, nv1l, n2)) if lo1 else None,
                     out(gderivs_max1, ("d1u",), (nb, nv1u, n2)) if hi1 else None, axis=1)
        for on, gd1, gd2, gml, gmu, nv2 in ((lo2, gd1min, gderivs_min2, gmixed_derivs_min1_min2,
                                             gmixed_derivs_max1_min2, nv2l),
                                            (hi2, gd1max, gderivs_max2, gmixed_derivs_min1_max2,
                                             gmixed_derivs_max1_max2, nv2u)):
            if not on:
                continue
            tag = "l" if gd1 is gd1min else "u"
            b1.build_vjp(gd1, out(gd2, ("d2" + tag,), (nb, n1, nv2)),
                         out(gml, ("ml" + tag,), (nb, nv1l, nv2)) if lo1 else None,
                         out(gmu, ("mu" + tag,), (nb, nv1u, nv2)) if hi1 else None, axis=1)
        return gvals


class SplineEvaluator2D:
    """ddc::SplineEvaluator2D (spline_evaluator_2d.hpp:37-826) on (batch, size1, size2) coefficients."""

    def __init__(self, ev1: SplineEvaluator, ev2: SplineEvaluator):
        self.ev1, self.ev2 = ev1, ev2

    # workspace of the packed evaluation: at most this many bytes are kept
    PACKED_WORKSPACE_MAX = 1 << 30

    def __call__(self, spline_eval, coords_eval, spline_coef, order1=0, order2=0, packed=None):
        """coords_eval: (batch, npts, 2); spline_eval: (batch, npts); spline_coef: (batch, s1, s2).
        packed: None picks the packed-coefficient path (ddcx_evaluator2d_eval_packed)
        when there are at least twice as many points as cells per batch entry;
        True / False forces it on / off."""
        _check_f64_cuda(coords_eval, "coords_eval")
        _check_f64_cuda(spline_eval, "spline_eval")
        _check_f64_cuda(spline_coef, "spline_coef")
        c = coords_eval.contiguous()
        squeeze = spline_coef.dim() == 2
        coef = spline_coef.unsqueeze(0) if squeeze else spline_coef
        nb = coef.shape[0]
        if c.shape[-1] != 2 or c.numel() % (2 * nb):
            raise ValueError("coords_eval must be (batch, npts, 2)")
        npts = c.numel() // (2 * nb)
        if not spline_eval.is_contiguous() or spline_eval.numel() != nb * npts:
            raise ValueError("spline_eval must be contiguous with one value per point")
        b1, b2 = self.ev1.bsplines, self.ev2.bsplines
        if coef.shape[1] != b1.size() or coef.shape[2] != b2.size():
            raise ValueError("spline_coef has the wrong number of coefficients")
        if packed is None:
            packed = npts >= 2 * b1.ncells() * b2.ncells()
        if packed and nb * npts > 0:
            per = C.c_int64(0)
            call("ddcx_evaluator2d_eval_workspace", self.ev1._h, self.ev2._h, 1, C.byref(per))
            entries = max(1, min(nb, self.PACKED_WORKSPACE_MAX // max(per.value, 1)))
            need = per.value * entries
            ws = getattr(self, "_pack_ws", None)
            if ws is None or ws.numel() < need or ws.device != c.device:
                ws = self._pack_ws = torch.empty(need, dtype=torch.uint8, device=c.device)
            call("ddcx_evaluator2d_eval_packed", self.ev1._h, self.ev2._h, int(order1), int(order2),
                 _ptr(spline_eval), _ptr(c), npts, _ptr(coef), coef.stride(0), coef.stride(1), coef.stride(2), nb,
                 _ptr(ws), ws.numel(), _stream())
            return spline_eval
        call("ddcx_evaluator2d_eval", self.ev1._h, self.ev2._h, int(order1), int(order2), _ptr(spline_eval), _ptr(c),
             npts, _ptr(coef), coef.stride(0), coef.stride(1), coef.stride(2), nb, _stream())
        return spline_eval

    def eval_vjp(self, coords_eval, w, spline_coef, gx, gcoef):
        """Adjoint of the value form: gcoef (overwritten, same shape as spline_coef)
        and gx (same shape as coords_eval) from the weights w (batch, npts)."""
        c = coords_eval.contiguous()
        ww = w.contiguous()
        squeeze = spline_coef.dim() == 2
        coef = spline_coef.unsqueeze(0) if squeeze else spline_coef
        gc = gcoef.unsqueeze(0) if squeeze else gcoef
        if tuple(gc.stride()) != tuple(coef.stride()):
            raise ValueError("gcoef must have the strides of spline_coef")
        if not gx.is_contiguous():
            raise ValueError("gx must be contiguous")
        nb = coef.shape[0]
        npts = c.numel() // (2 * nb)
        need = C.c_int64(0)
        call("ddcx_evaluator2d_eval_vjp_workspace", self.ev1._h, self.ev2._h, npts, nb, C.byref(need))
        if need.value > 0:
            # points binned by cell tile (eval2d_binned.hip); the workspace is
            # kept for the next call of the same or a smaller size
            ws = getattr(self, "_work", None)
            if ws is None or ws.numel() < need.value or ws.device != c.device:
                ws = self._work = torch.empty(need.value, dtype=torch.uint8, device=c.device)
            call("ddcx_evaluator2d_eval_vjp_binned", self.ev1._h, self.ev2._h, _ptr(c), _ptr(ww), _ptr(gx), npts,
                 _ptr(coef), _ptr(gc), coef.stride(0), coef.stride(1), coef.stride(2), nb, _ptr(ws), ws.numel(),
                 _stream())
        else:
            call("ddcx_evaluator2d_eval_vjp", self.ev1._h, self.ev2._h, _ptr(c), _ptr(ww), _ptr(gx), npts,
                 _ptr(coef), _ptr(gc), coef.stride(0), coef.stride(1), coef.stride(2), nb, _stream())
        return gx, gcoef

    def deriv_dim_1(self, spline_eval, coords_eval, spline_coef):
        return self(spline_eval, coords_eval, spline_coef, 1, 0)

    def deriv_dim_2(self, spline_eval, coords_eval, spline_coef):
        return self(spline_eval, coords_eval, spline_coef, 0, 1)

    def deriv_1_and_2(self, spline_eval, coords_eval, spline_coef):
        return self(spline_eval, coords_eval, spline_coef, 1, 1)


class SplineBuilder3D:
    """ddc::SplineBuilder3D (spline_builder_3d.hpp:45-1090): arrays are
    (batch, n1, n2, n3) (or (n1, n2, n3)).  With Hermite closures the call
    takes the reference's derivative arguments by name (derivs_min1 ...
    mixed_derivs_max1_max2_max3, shaped as the reference's batched derivative
    domains, e.g. derivs_min2 (batch, n1, nbe2, n3), mixed_derivs_min1_max3
    (batch, nbe1, n2, nbe3)).  As the reference does: the dimension-2 and
    dimension-3 derivatives and the 2-3 cross derivatives are built along
    dimension 1 (with the mixed derivatives involving dimension 1 as their
    Hermite data), the values along dimension 1, then SplineBuilder2D over
    dimensions 2 and 3.  Intermediates are torch tensors owned by this
    object."""

    DERIV_NAMES = tuple(
        [f"derivs_{s}{k}" for k in (1, 2, 3) for s in ("min", "max")]
        + [f"mixed_derivs_{a}{p}_{b}{q}" for p, q in ((1, 2), (2, 3), (1, 3)) for b in ("min", "max")
           for a in ("min", "max")]
        + [f"mixed_derivs_{a}1_{b}2_{c}3" for c in ("min", "max") for b in ("min", "max") for a in ("min", "max")])

    def __init__(self, builder1: SplineBuilder, builder2: SplineBuilder, builder3: SplineBuilder):
        self.builders = (builder1, builder2, builder3)
        self._b23 = SplineBuilder2D(builder2, builder3)
        self._tmp = {}

    def _scratch(self, key, shape, device):
        t = self._tmp.get(key)
        if t is None or tuple(t.shape) != tuple(shape) or t.device != device:
            t = self._tmp[key] = torch.empty(shape, dtype=torch.float64, device=device)
        return t

    def __call__(self, spline, vals, **derivs):
        """SplineBuilder3D::operator() (spline_builder_3d.hpp:779-1090)."""
        unknown = set(derivs) - set(self.DERIV_NAMES)
        if unknown:
            raise TypeError(f"unknown derivative arguments: {sorted(unknown)}")
        squeeze = vals.dim() == 3

        def b(t):
            return None if t is None else (t.unsqueeze(0) if squeeze else t)

        v, s = b(vals), b(spline)
        d = {k: b(t) for k, t in derivs.items() if t is not None}
        nb = v.shape[0]
        b1, b2, b3 = self.builders
        H = SplineBuilderClosure.HERMITE
        s1 = b1.bsplines.size()
        dev = v.device

        def along1(key, name, lo, hi):
            if name not in d:
                raise ValueError(f"{name} is required by the Hermite closures")
            src = d[name]
            out = self._scratch(key, (nb, s1) + tuple(src.shape[2:]), dev)
            b1(out, src, d.get(lo), d.get(hi), axis=1)
            return out

        d2 = {sd: along1(("d2", sd), f"derivs_{sd}2", f"mixed_derivs_min1_{sd}2", f"mixed_derivs_max1_{sd}2")
              for sd, bc in (("min", b2.lower), ("max", b2.upper)) if bc == H}
        d3 = {sd: along1(("d3", sd), f"derivs_{sd}3", f"mixed_derivs_min1_{sd}3", f"mixed_derivs_max1_{sd}3")
              for sd, bc in (("min", b3.lower), ("max", b3.upper)) if bc == H}
        m23 = {}
        for a2, bc2 in (("min", b2.lower), ("max", b2.upper)):
            for a3, bc3 in (("min", b3.lower), ("max", b3.upper)):
                if bc2 == H and bc3 == H:
                    m23[a2, a3] = along1(("m23", a2, a3), f"mixed_derivs_{a2}2_{a3}3",
                                         f"mixed_derivs_min1_{a2}2_{a3}3", f"mixed_derivs_max1_{a2}2_{a3}3")
        spline1 = self._scratch("s1", (nb, s1) + tuple(v.shape[2:]), dev)
        b1(spline1, v, d.get("derivs_min1"), d.get("derivs_max1"), axis=1)

        def flat(t):
            return None if t is None else t.reshape((nb * s1,) + tuple(t.shape[2:]))

        self._b23(flat(s), flat(spline1), flat(d2.get("min")), flat(d2.get("max")), flat(d3.get("min")),
                  flat(d3.get("max")), flat(m23.get(("min", "min"))), flat(m23.get(("max", "min"))),
                  flat(m23.get(("min", "max"))), flat(m23.get(("max", "max"))))
        return spline

    def build_vjp(self, gspline, gvals, **gderivs):
        """Adjoint of __call__: gvals (batch, n1, n2, n3) from gspline (batch,
        size1, size2, size3) and, with Hermite closures, the gradients of the
        derivative inputs, passed by the argument's name with a "g" prefix
        (gderivs_min1, gmixed_derivs_min1_max2_min3, ...; each shaped as the
        __call__ argument, overwritten; omitted: not returned).  __call__'s
        passes in reverse: SplineBuilder2D.build_vjp over dimensions 2 and 3
        gives the gradients of the dimension-1 splines (of the values, of the
        dimension-2 and -3 derivatives and of the 2-3 cross derivatives), then
        transposed dimension-1 solves give the gradients of the inputs."""
        unknown = set(gderivs) - {"g" + n for n in self.DERIV_NAMES}
        if unknown:
            raise TypeError(f"unknown gradient arguments: {sorted(unknown)}")
        squeeze = gspline.dim() == 3

        def b(t):
            return None if t is None else (t.unsqueeze(0) if squeeze else t)

        g, gv = b(gspline), b(gvals)
        _check_f64_cuda(gv, "gvals")
        nb, n1, n2, n3 = gv.shape
        b1, b2, b3 = self.builders
        s1 = b1.bsplines.size()
        H = SplineBuilderClosure.HERMITE
        dev = g.device
        nv = {(k, "min"): bb.s_nbv_xmin for k, bb in zip((1, 2, 3), self.builders)}
        nv.update({(k, "max"): bb.s_nbv_xmax for k, bb in zip((1, 2, 3), self.builders)})
        herm = {(k, sd): (bb.lower if sd == "min" else bb.upper) == H
                for k, bb in zip((1, 2, 3), self.builders) for sd in ("min", "max")}

        def out(name, shape):
            # the caller's gradient, or scratch when it is not asked for
            t = gderivs.get("g" + name)
            return b(t) if t is not None else self._scratch(("g", name), shape, dev)

        def scratch(key, shape):
            return self._scratch(("g",) + key, shape, dev)

        def flat(t):
            return None if t is None else t.reshape((nb * s1,) + tuple(t.shape[2:]))

        sides = ("min", "max")
        gs1 = scratch(("s1",), (nb, s1, n2, n3))
        gd2 = {sd: scratch(("d2", sd), (nb, s1, nv[2, sd], n3)) for sd in sides if herm[2, sd]}
        gd3 = {sd: scratch(("d3", sd), (nb, s1, n2, nv[3, sd])) for sd in sides if herm[3, sd]}
        gm23 = {(a2, a3): scratch(("m23", a2, a3), (nb, s1, nv[2, a2], nv[3, a3]))
                for a2 in sides for a3 in sides if herm[2, a2] and herm[3, a3]}
        self._b23.build_vjp(flat(g), flat(gs1), flat(gd2.get("min")), flat(gd2.get("max")), flat(gd3.get("min")),
                            flat(gd3.get("max")), flat(gm23.get(("min", "min"))), flat(gm23.get(("max", "min"))),
                            flat(gm23.get(("min", "max"))), flat(gm23.get(("max", "max"))))

        def along1(src, name, lo, hi, rest):
            # transposed dimension-1 solve of one of __call__'s dimension-1 builds
            b1.build_vjp(src, out(name, (nb, n1) + rest),
                         out(lo, (nb, nv[1, "min"]) + rest) if herm[1, "min"] else None,
                         out(hi, (nb, nv[1, "max"]) + rest) if herm[1, "max"] else None, axis=1)

        b1.build_vjp(gs1, gv, out("derivs_min1", (nb, nv[1, "min"], n2, n3)) if herm[1, "min"] else None,
                     out("derivs_max1", (nb, nv[1, "max"], n2, n3)) if herm[1, "max"] else None, axis=1)
        for sd, t in gd2.items():
            along1(t, f"derivs_{sd}2", f"mixed_derivs_min1_{sd}2", f"mixed_derivs_max1_{sd}2", (nv[2, sd], n3))
        for sd, t in gd3.items():
            along1(t, f"derivs_{sd}3", f"mixed_derivs_min1_{sd}3", f"mixed_derivs_max1_{sd}3", (n2, nv[3, sd]))
        for (a2, a3), t in gm23.items():
            along1(t, f"mixed_derivs_{a2}2_{a3}3", f"mixed_derivs_min1_{a2}2_{a3}3",
                   f"mixed_derivs_max1_{a2}2_{a3}3", (nv[2, a2], nv[3, a3]))
        return gvals


class SplineEvaluatorND:
    """ddc::SplineEvaluatorND (spline_evaluator_nd.hpp:49-827) over one 1D
    evaluator per dimension (1 to 4 dimensions; Null or Periodic rules).
    Coefficients are (batch, size_1, ..., size_N) with any strides, points
    (batch, npts, N) and values (batch, npts)."""

    def __init__(self, *evaluators: SplineEvaluator):
        if not 1 <= len(evaluators) <= 4:
            raise ValueError("SplineEvaluatorND takes 1 to 4 evaluators")
        self.evaluators = tuple(evaluators)
        self._hs = (C.c_void_p * len(evaluators))(*[e._h for e in evaluators])

    def _coef(self, spline_coef):
        nd = len(self.evaluators)
        coef = spline_coef.unsqueeze(0) if spline_coef.dim() == nd else spline_coef
        if coef.dim() != nd + 1:
            raise ValueError(f"spline_coef must have {nd} (or {nd} + 1 batched) dimensions")
        st = (C.c_int64 * nd)(*[coef.stride(k + 1) for k in range(nd)])
        return coef, st

    def deriv(self, orders, spline_eval, coords_eval, spline_coef, binned=None):
        """SplineEvaluatorND::deriv (spline_evaluator_nd.hpp:445-590): orders
        is one derivative order per dimension.  binned: None takes the
        binned-by-tile evaluation (ddcx_evaluatornd_eval_binned, the same
        values bit for bit) when the grid allows it and a batch entry has at
        least one point per eight cells; True requires it; False never."""
        nd = len(self.evaluators)
        coef, st = self._coef(spline_coef)
        _check_f64_cuda(coords_eval, "coords_eval")
        _check_f64_cuda(spline_eval, "spline_eval")
        c = coords_eval.contiguous()
        if c.shape[-1] != nd:
            raise ValueError(f"coords_eval must end with a dimension of size {nd}")
        if not spline_eval.is_contiguous():
            raise ValueError("spline_eval must be contiguous")
        nb = coef.shape[0]
        npts = c.numel() // (nd * nb)
        if spline_eval.numel() != npts * nb:
            raise ValueError("spline_eval must hold one value per point")
        ords = None if orders is None else (C.c_int * nd)(*[int(o) for o in orders])
        need = C.c_int64(0)
        if binned is not False:
            call("ddcx_evaluatornd_eval_workspace", self._hs, nd, npts, nb, C.byref(need))
            if binned is True and need.value == 0:
                raise NotImplementedError("the binned evaluation takes three dimensions of one degree <= 5")
            if binned is None:
                cells = int(np.prod([e.bsplines.ncells() for e in self.evaluators]))
                if 8 * npts < cells:
                    need.value = 0
        if need.value > 0:
            ws = getattr(self, "_work", None)
            if ws is None or ws.numel() < need.value or ws.device != c.device:
                ws = self._work = torch.empty(need.value, dtype=torch.uint8, device=c.device)
            call("ddcx_evaluatornd_eval_binned", self._hs, nd, ords, _ptr(spline_eval), _ptr(c), npts, _ptr(coef),
                 coef.stride(0), st, nb, _ptr(ws), ws.numel(), _stream())
        else:
            call("ddcx_evaluatornd_eval", self._hs, nd, ords, _ptr(spline_eval), _ptr(c), npts, _ptr(coef),
                 coef.stride(0), st, nb, _stream())
        return spline_eval

    def __call__(self, spline_eval, coords_eval, spline_coef, binned=None):
        """SplineEvaluatorND::operator() (spline_evaluator_nd.hpp:300-440)."""
        return self.deriv(None, spline_eval, coords_eval, spline_coef, binned=binned)

    def integrate(self, integrals_out, spline_coef):
        """SplineEvaluatorND::integrate (spline_evaluator_nd.hpp:597-632):
        integrals_out (batch,) receives the integral over the whole domain."""
        nd = len(self.evaluators)
        coef, st = self._coef(spline_coef)
        nb = coef.shape[0]
        out = integrals_out.reshape(-1)
        if out.numel() != nb:
            raise ValueError("integrals_out must hold one value per batch entry")
        # the partial sums of the split reduction go to a device workspace
        need = C.c_int64()
        call("ddcx_evaluatornd_integrate_workspace", self._hs, nd, nb, C.byref(need))
        ws = torch.empty(max(need.value, 1), dtype=torch.uint8, device=coef.device)
        call("ddcx_evaluatornd_integrate_ws", self._hs, nd, _ptr(out), out.stride(0), _ptr(coef), coef.stride(0), st,
             nb, _ptr(ws), need.value, _stream())
        return integrals_out


class SplineEvaluator3D(SplineEvaluatorND):
    """ddc::SplineEvaluator3D (spline_evaluator_3d.hpp:56-1019): the
    three-dimensional SplineEvaluatorND."""

    def __init__(self, ev1: SplineEvaluator, ev2: SplineEvaluator, ev3: SplineEvaluator):
        super().__init__(ev1, ev2, ev3)

    def deriv_dim_1(self, spline_eval, coords_eval, spline_coef):
        return self.deriv((1, 0, 0), spline_eval, coords_eval, spline_coef)

    def deriv_dim_2(self, spline_eval, coords_eval, spline_coef):
        return self.deriv((0, 1, 0), spline_eval, coords_eval, spline_coef)

    def deriv_dim_3(self, spline_eval, coords_eval, spline_coef):
        return self.deriv((0, 0, 1), spline_eval, coords_eval, spline_coef)

    def deriv_1_and_2(self, spline_eval, coords_eval, spline_coef):
        return self.deriv((1, 1, 0), spline_eval, coords_eval, spline_coef)

    def deriv_1_and_3(self, spline_eval, coords_eval, spline_coef):
        return self.deriv((1, 0, 1), spline_eval, coords_eval, spline_coef)

    def deriv_2_and_3(self, spline_eval, coords_eval, spline_coef):
        return self.deriv((0, 1, 1), spline_eval, coords_eval, spline_coef)

    def deriv_1_2_and_3(self, spline_eval, coords_eval, spline_coef):
        return self.deriv((1, 1, 1), spline_eval, coords_eval, spline_coef)
