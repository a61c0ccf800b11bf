"""ddc_amd: MI355X (gfx950) implementation of DDC's batched spline hot path.

Native code lives in ``libddcx.so`` (C++20 host setup + hand-written HIP
kernels, C ABI in ``include/ddcx.h``); this package is the host-side mirror of
the reference's ``ddc::SplineBuilder`` / ``ddc::SplineEvaluator`` interface.
"""
from ._lib import LIB_PATH, DDCXError, lib  # noqa: F401
from . import autograd  # noqa: F401
from .splines import (  # noqa: F401
    ConstantExtrapolationRule,
    NonUniformBSplines,
    NullExtrapolationRule,
    PeriodicExtrapolationRule,
    Precision,
    SplineBuilder,
    SplineBuilder2D,
    SplineBuilder3D,
    SplineBuilderClosure,
    SplineEvaluator,
    SplineEvaluator2D,
    SplineEvaluator3D,
    SplineEvaluatorND,
    SplineSolver,
    UniformBSplines,
    advect,
    batched_layout,
    build_eval,
    integrals,
    knots_as_interpolation_points,
    n_boundary_equations,
)

__version__ = "0.1.0"
