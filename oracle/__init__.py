"""CPU oracle for DDC's batched spline path (test checker and CPU baseline).

This package is test infrastructure.  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it, and only to check or
time the product; the product library never links to or calls it.

``liboracle.so`` is built from ``ddc_oracle.c`` (a step-by-step FP64
restatement of the reference algorithm, compiled with ``-ffp-contract=off``)
by ``oracle/Makefile``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

PERIODIC, HERMITE, HOMOGENEOUS_HERMITE, GREVILLE = 0, 1, 2, 3

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_vp = C.c_void_p
_L = C.c_long


def build() -> str:
    """Compile liboracle.so in place (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.or_last_error.restype = C.c_char_p
        L.or_problem_create.restype = _vp
        L.or_problem_create.argtypes = [C.c_int] * 7
        L.or_problem_set.argtypes = [_vp, C.c_int, C.c_int, C.c_double]
        L.or_problem_get.argtypes = [_vp, C.c_int, C.c_int]
        L.or_problem_get.restype = C.c_double
        L.or_problem_setup.argtypes = [_vp]
        L.or_problem_type.argtypes = [_vp]
        L.or_problem_rhs_rows.argtypes = [_vp]
        L.or_problem_solve.argtypes = [_vp, _dp, C.c_int, C.c_int]
        L.or_problem_destroy.argtypes = [_vp]
        L.or_problem_q.argtypes = [_vp]
        L.or_problem_q.restype = C.POINTER(C.c_double)
        L.or_problem_ipiv.argtypes = [_vp]
        L.or_problem_ipiv.restype = C.POINTER(C.c_int)
        L.or_spline_create.restype = _vp
        L.or_spline_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                       _vp, _vp, C.c_int]
        for name in ["or_spline_npts", "or_spline_nbasis", "or_spline_size", "or_spline_offset",
                     "or_spline_matrix_type", "or_spline_nknots"]:
            getattr(L, name).argtypes = [_vp]
        for name in ["or_spline_dx", "or_spline_rmin", "or_spline_rmax"]:
            getattr(L, name).argtypes = [_vp]
            getattr(L, name).restype = C.c_double
        L.or_spline_problem.argtypes = [_vp]
        L.or_spline_problem.restype = _vp
        L.or_spline_points.argtypes = [_vp, _dp]
        L.or_spline_knots.argtypes = [_vp, _dp]
        L.or_spline_integrals.argtypes = [_vp, _dp]
        L.or_spline_eval_basis.argtypes = [_vp, C.c_double, _dp]
        L.or_spline_eval_basis_and_n_derivs.argtypes = [_vp, C.c_double, C.c_int, _dp]
        L.or_spline_build.argtypes = [_vp, _dp, _L, _L, _L, _L, _dp, _L, _L, _L, _L, _vp, _vp, _L, _L, _L, _L, _L]
        L.or_spline_build_vjp.argtypes = [_vp, _dp, _L, _L, _L, _L, _dp, _L, _L, _L, _L, _vp, _vp,
                                          _L, _L, _L, _L, _L]
        L.or_spline_eval.argtypes = [_vp, _dp, _dp, _L, _L, _L, _L, _L, _dp, _L, _L, _L, _L, _L, C.c_int,
                                     C.c_int, C.c_int, _vp]
        L.or_spline_eval_vjp.argtypes = [_vp, _dp, _dp, _dp, _L, _L, _L, _L, _L, _dp, _dp, _L, _L, _L, _L, _L,
                                         C.c_int, C.c_int, _vp]
        L.or_spline_integrate.argtypes = [_vp, _dp, _L, _L, _L, _L, _dp, _L, _L]
        L.or_spline_quadrature.argtypes = [_vp, _dp]
        L.or_spline2d_eval.argtypes = [_vp, _vp, _dp, _dp, _L, _dp, _L, _L, _L, _L, C.c_int, C.c_int,
                                       np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS"), _vp]
        L.or_splinend_eval.argtypes = [C.POINTER(_vp), C.c_int, _dp, _dp, _L, _dp, _L, C.POINTER(_L), _L,
                                       C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.or_splinend_integrate.argtypes = [C.POINTER(_vp), C.c_int, _dp, _dp, _L, C.POINTER(_L), _L]
        L.or_spline_destroy.argtypes = [_vp]
        _LIB = L
    return _LIB


def _err():
    return lib().or_last_error().decode()


class Problem:
    """A SplinesLinearProblem restated on the CPU (for solver known-answer tests)."""

    KINDS = {"dense": 0, "band": 1, "block": 2, "periodic": 3, "dense2x2": 4, "dense3x3": 5, "band_storage": 6}
    TYPES = ["dense", "band", "pds_band", "pds_tridiag", "2x2", "3x3"]

    def __init__(self, kind, n, kl=0, ku=0, pds=False, bottom=0, top=0):
        self.h = lib().or_problem_create(self.KINDS[kind], n, kl, ku, int(pds), bottom, top)
        self.n = n

    def __del__(self):
        if getattr(self, "h", None):
            try:
                lib().or_problem_destroy(self.h)
            except Exception:
                pass
            self.h = None

    def set(self, i, j, v):
        lib().or_problem_set(self.h, i, j, float(v))

    def get(self, i, j):
        return lib().or_problem_get(self.h, i, j)

    @property
    def type(self):
        return self.TYPES[lib().or_problem_type(self.h)]

    @property
    def rhs_rows(self):
        return lib().or_problem_rhs_rows(self.h)

    def setup(self):
        info = lib().or_problem_setup(self.h)
        if info:
            raise RuntimeError(f"factorisation failed with error code {info}")

    def solve(self, b, transpose=False):
        b = np.ascontiguousarray(b, dtype=np.float64)
        assert b.shape[0] == self.rhs_rows
        lib().or_problem_solve(self.h, b, b.shape[1], int(transpose))
        return b


class Spline:
    """A 1D spline builder + evaluator (SplineBuilder/SplineEvaluator) on the CPU."""

    def __init__(self, degree, uniform, bc_lower, bc_upper, ncells=None, rmin=0.0, rmax=1.0, breaks=None,
                 points=None):
        self._keep = []
        bp = None
        if not uniform:
            br = np.ascontiguousarray(breaks, dtype=np.float64)
            self._keep.append(br)
            ncells = br.size - 1
            bp = br.ctypes.data
        pp, npts = None, 0
        if points is not None:
            pa = np.ascontiguousarray(points, dtype=np.float64)
            self._keep.append(pa)
            pp, npts = pa.ctypes.data, pa.size
        self.h = lib().or_spline_create(degree, int(uniform), bc_lower, bc_upper, ncells, rmin, rmax, bp, pp, npts)
        if not self.h:
            raise RuntimeError(_err())
        self.degree, self.uniform, self.bc_lower, self.bc_upper = degree, uniform, bc_lower, bc_upper
        self.periodic = bc_lower == PERIODIC
        self.ncells = ncells
        L = lib()
        self.npts = L.or_spline_npts(self.h)
        self.nbasis = L.or_spline_nbasis(self.h)
        self.size = L.or_spline_size(self.h)
        self.offset = L.or_spline_offset(self.h)
        self.dx = L.or_spline_dx(self.h)
        self.rmin = L.or_spline_rmin(self.h)
        self.rmax = L.or_spline_rmax(self.h)
        self.matrix_type = Problem.TYPES[L.or_spline_matrix_type(self.h)]
        nbe = lambda bc: degree // 2 if bc in (HERMITE, HOMOGENEOUS_HERMITE) else 0
        self.nbe_l, self.nbe_r = nbe(bc_lower), nbe(bc_upper)
        self.nbv_l = 0 if bc_lower == HOMOGENEOUS_HERMITE else self.nbe_l
        self.nbv_r = 0 if bc_upper == HOMOGENEOUS_HERMITE else self.nbe_r

    def __del__(self):
        if getattr(self, "h", None):
            try:
                lib().or_spline_destroy(self.h)
            except Exception:
                pass
            self.h = None

    def points(self):
        out = np.empty(self.npts)
        lib().or_spline_points(self.h, out)
        return out

    def knots(self):
        out = np.empty(lib().or_spline_nknots(self.h))
        lib().or_spline_knots(self.h, out)
        return out

    def integrals(self):
        out = np.empty(self.size)
        lib().or_spline_integrals(self.h, out)
        return out

    def eval_basis(self, x):
        v = np.empty(self.degree + 1)
        j = lib().or_spline_eval_basis(self.h, float(x), v)
        return j, v

    def eval_basis_and_n_derivs(self, x, n):
        d = np.empty((self.degree + 1) * (n + 1))
        j = lib().or_spline_eval_basis_and_n_derivs(self.h, float(x), n, d)
        return j, d.reshape(self.degree + 1, n + 1)

    def matrix(self):
        p = lib().or_spline_problem(self.h)
        n = self.nbasis
        # the matrix is factored in place by setup; this returns the factored storage view
        return p

    @staticmethod
    def _lay(a, axis_interp):
        """Strides (in elements) of a 2-D C-contiguous array (rows along axis_interp)."""
        s = [st // 8 for st in a.strides]
        if axis_interp == 0:  # [interp][batch]
            return s[0], a.shape[1], s[1], 0
        return s[1], a.shape[0], s[0], 0

    def build(self, vals, derivs_l=None, derivs_r=None, layout="batch_contiguous"):
        """vals: (npts, B) for batch_contiguous or (B, npts) for interp_contiguous.
        Returns coef in the same layout, with `size` coefficients per line."""
        ax = 0 if layout == "batch_contiguous" else 1
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        B = vals.shape[1 - ax]
        coef = np.empty((self.size, B) if ax == 0 else (B, self.size))
        vl, cl = self._lay(vals, ax), self._lay(coef, ax)
        dlp = drp = None
        dlay = (1, B, 0, 0)
        if derivs_l is not None and self.nbv_l:
            dl = np.ascontiguousarray(derivs_l, dtype=np.float64)
            self._tmp_dl = dl
            dlp = dl.ctypes.data
            dlay = self._lay(dl, ax)
        if derivs_r is not None and self.nbv_r:
            dr = np.ascontiguousarray(derivs_r, dtype=np.float64)
            self._tmp_dr = dr
            drp = dr.ctypes.data
            dlay = self._lay(dr, ax)
        if (self.bc_lower == HERMITE and dlp is None) or (self.bc_upper == HERMITE and drp is None):
            raise ValueError("Hermite closure needs boundary derivatives")
        lib().or_spline_build(self.h, vals, *vl, coef, *cl, dlp, drp, *dlay, B)
        return coef

    def build_vjp(self, gcoef, layout="batch_contiguous"):
        ax = 0 if layout == "batch_contiguous" else 1
        gcoef = np.ascontiguousarray(gcoef, dtype=np.float64)
        B = gcoef.shape[1 - ax]
        gvals = np.empty((self.npts, B) if ax == 0 else (B, self.npts))
        gdl = np.zeros((max(self.nbv_l, 1), B) if ax == 0 else (B, max(self.nbv_l, 1)))
        gdr = np.zeros((max(self.nbv_r, 1), B) if ax == 0 else (B, max(self.nbv_r, 1)))
        lib().or_spline_build_vjp(self.h, gcoef, *self._lay(gcoef, ax), gvals, *self._lay(gvals, ax),
                                  gdl.ctypes.data, gdr.ctypes.data, *self._lay(gdl, ax), B)
        if ax == 0:
            return gvals, gdl[: self.nbv_l], gdr[: self.nbv_r]
        return gvals, gdl[:, : self.nbv_l], gdr[:, : self.nbv_r]

    @staticmethod
    def _rule_pos(eval_pos):
        """(lower, upper) eval_pos of constant rules, or None for the bounds."""
        if eval_pos is None:
            return None, None
        rp = np.ascontiguousarray(eval_pos, dtype=np.float64)
        assert rp.shape == (2,)
        return rp, rp.ctypes.data

    def eval(self, coords, coef, deriv=0, layout="batch_contiguous", extrap=("null", "null"), eval_pos=None):
        """eval_pos: (lower, upper) positions of constant rules (None: rmin, rmax)."""
        ax = 0 if layout == "batch_contiguous" else 1
        keep, rp = self._rule_pos(eval_pos)
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        coef = np.ascontiguousarray(coef, dtype=np.float64)
        B = coords.shape[1 - ax]
        P = coords.shape[ax]
        out = np.empty_like(coords)
        rules = [0 if e == "null" else 1 for e in extrap]
        lib().or_spline_eval(self.h, coords, out, *self._lay(coords, ax), P, coef, *self._lay(coef, ax), B,
                             deriv, rules[0], rules[1], rp)
        del keep
        return out

    def eval_vjp(self, coords, w, coef, layout="batch_contiguous", extrap=("null", "null"), eval_pos=None):
        ax = 0 if layout == "batch_contiguous" else 1
        keep, rp = self._rule_pos(eval_pos)
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        w = np.ascontiguousarray(w, dtype=np.float64)
        coef = np.ascontiguousarray(coef, dtype=np.float64)
        B = coords.shape[1 - ax]
        P = coords.shape[ax]
        gx = np.empty_like(coords)
        gc = np.empty_like(coef)
        rules = [0 if e == "null" else 1 for e in extrap]
        lib().or_spline_eval_vjp(self.h, coords, w, gx, *self._lay(coords, ax), P, coef, gc,
                                 *self._lay(coef, ax), B, rules[0], rules[1], rp)
        del keep
        return gc, gx

    def integrate(self, coef, layout="batch_contiguous"):
        ax = 0 if layout == "batch_contiguous" else 1
        coef = np.ascontiguousarray(coef, dtype=np.float64)
        B = coef.shape[1 - ax]
        out = np.empty(B)
        lib().or_spline_integrate(self.h, coef, *self._lay(coef, ax), out, 1, B)
        return out

    def quadrature_coefficients(self):
        out = np.empty(self.nbv_l + (self.nbasis - self.nbv_l - self.nbv_r) + self.nbv_r)
        lib().or_spline_quadrature(self.h, out)
        return (out[: self.nbv_l], out[self.nbv_l: out.size - self.nbv_r], out[out.size - self.nbv_r:])


def spline2d_build(s1: Spline, s2: Spline, vals, derivs_min1=None, derivs_max1=None, derivs_min2=None,
                   derivs_max2=None, mixed_min1_min2=None, mixed_max1_min2=None, mixed_min1_max2=None,
                   mixed_max1_max2=None):
    """SplineBuilder2D::operator() (spline_builder_2d.hpp:514-651):
    vals (nbatch, n1, n2) -> coef (nbatch, size1, size2).  Hermite data, as
    in the reference: derivs_*1 (nbatch, nbe1, n2) along dimension 1 at each
    dimension-2 point, derivs_*2 (nbatch, n1, nbe2), mixed_* (nbatch, nbe1,
    nbe2) at the corners.  Dimension-2 derivatives are first built along
    dimension 1 (with the corner derivatives as their dimension-1 Hermite
    data, :583-640), then dimension 2 is built over the dimension-1 spline
    with those as its Hermite data (:641-650)."""
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    nb, n1, n2 = vals.shape
    L = lib()

    def c(a):
        return None if a is None else np.ascontiguousarray(a, dtype=np.float64)

    def build1(src, n_cols, dmin, dmax):
        # builder1 along dim 1 of (nb, n1, n_cols), batched over (batch, column)
        out = np.empty((nb, s1.size, n_cols))
        dmin, dmax = c(dmin), c(dmax)
        nbe = (dmin if dmin is not None else dmax).shape[1] if (dmin is not None or dmax is not None) else 1
        L.or_spline_build(s1.h, src, n_cols, n_cols, 1, n1 * n_cols, out, n_cols, n_cols, 1, s1.size * n_cols,
                          None if dmin is None else dmin.ctypes.data, None if dmax is None else dmax.ctypes.data,
                          n_cols, n_cols, 1, nbe * n_cols, nb * n_cols)
        return out

    d1min = d1max = None
    if derivs_min2 is not None:
        d1min = build1(c(derivs_min2), c(derivs_min2).shape[2], mixed_min1_min2, mixed_max1_min2)
    spline1 = build1(vals, n2, derivs_min1, derivs_max1)
    if derivs_max2 is not None:
        d1max = build1(c(derivs_max2), c(derivs_max2).shape[2], mixed_min1_max2, mixed_max1_max2)
    coef = np.empty((nb, s1.size, s2.size))
    nbe2 = (d1min if d1min is not None else d1max).shape[2] if (d1min is not None or d1max is not None) else 1
    # builder2 along dim 2 batched over (batch, bsplines1)
    L.or_spline_build(s2.h, spline1, 1, s1.size, n2, s1.size * n2, coef, 1, s1.size, s2.size,
                      s1.size * s2.size, None if d1min is None else d1min.ctypes.data,
                      None if d1max is None else d1max.ctypes.data, 1, s1.size, nbe2, s1.size * nbe2, nb * s1.size)
    return coef


def spline2d_eval(s1: Spline, s2: Spline, coords, coef, d1=0, d2=0, rules=(0, 0, 0, 0), rule_pos=None):
    """coords (nbatch, npts, 2), coef (nbatch, size1, size2) -> (nbatch, npts).
    rules: lower 1, upper 1, lower 2, upper 2 (0 null, 1 constant).  rule_pos:
    None (constant rules at the guarded bound, clamping the other coordinate
    to its domain) or (pos[4], min[4], max[4]) per rule."""
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    coef = np.ascontiguousarray(coef, dtype=np.float64)
    nb, P, _ = coords.shape
    out = np.empty((nb, P))
    rp = None if rule_pos is None else np.ascontiguousarray(np.asarray(rule_pos, dtype=np.float64).reshape(12))
    lib().or_spline2d_eval(s1.h, s2.h, coords, out, P, coef, s1.size * s2.size, s2.size, 1, nb, d1, d2,
                           np.asarray(rules, dtype=np.int32), None if rp is None else rp.ctypes.data)
    return out


def splinend_build(splines, vals):
    """SplineBuilder3D::operator() (spline_builder_3d.hpp; SplineBuilder2D for
    two dimensions) without Hermite closures: one batched 1D build per
    dimension, dimension 1 first.  vals (nbatch, n_1, ..., n_N) ->
    coef (nbatch, size_1, ..., size_N)."""
    cur = np.ascontiguousarray(vals, dtype=np.float64)
    for k, s in enumerate(splines):
        ax = k + 1
        shape = list(cur.shape)
        outer = int(np.prod(shape[:ax]))
        inner = int(np.prod(shape[ax + 1:]))
        shape[ax] = s.size
        nxt = np.empty(shape)
        # rows along axis ax; lines enumerate (outer, inner) in row-major order
        lib().or_spline_build(s.h, cur, inner, inner, 1, cur.shape[ax] * inner, nxt, inner, inner, 1,
                              s.size * inner, None, None, 1, 1, 0, 0, outer * inner)
        cur = nxt
    return cur


def _build_axis(s: Spline, src, ax, dmin=None, dmax=None):
    """SplineBuilder::operator() along axis `ax` of a C-contiguous array,
    batched over every other axis (lines in row-major order).  dmin / dmax:
    the Hermite data, arrays shaped like src with nbv rows along `ax`."""
    src = np.ascontiguousarray(src, dtype=np.float64)
    shape = list(src.shape)
    outer = int(np.prod(shape[:ax]))
    inner = int(np.prod(shape[ax + 1:]))
    shape[ax] = s.size
    out = np.empty(shape)
    dmin = None if dmin is None or s.nbv_l == 0 else np.ascontiguousarray(dmin, dtype=np.float64)
    dmax = None if dmax is None or s.nbv_r == 0 else np.ascontiguousarray(dmax, dtype=np.float64)
    if (s.bc_lower == HERMITE and s.nbv_l and dmin is None) or (s.bc_upper == HERMITE and s.nbv_r and dmax is None):
        raise ValueError("Hermite closure needs boundary derivatives")
    nbe = (dmin if dmin is not None else dmax).shape[ax] if (dmin is not None or dmax is not None) else 1
    lib().or_spline_build(s.h, src, inner, inner, 1, src.shape[ax] * inner, out, inner, inner, 1, s.size * inner,
                          None if dmin is None else dmin.ctypes.data, None if dmax is None else dmax.ctypes.data,
                          inner, inner, 1, nbe * inner, outer * inner)
    return out


def spline3d_build(s1: Spline, s2: Spline, s3: Spline, vals, derivs=None):
    """SplineBuilder3D::operator() (spline_builder_3d.hpp:889-1090), with the
    reference's Hermite data: vals (nbatch, n1, n2, n3) -> coef (nbatch,
    size1, size2, size3).  derivs maps the reference's argument names to
    arrays: derivs_min1/max1 (nbatch, nbe1, n2, n3), derivs_min2/max2
    (nbatch, n1, nbe2, n3), derivs_min3/max3 (nbatch, n1, n2, nbe3),
    mixed_derivs_<a>1_<b>2 (nbatch, nbe1, nbe2, n3), mixed_derivs_<b>2_<c>3
    (nbatch, n1, nbe2, nbe3), mixed_derivs_<a>1_<c>3 (nbatch, nbe1, n2,
    nbe3) and mixed_derivs_<a>1_<b>2_<c>3 (nbatch, nbe1, nbe2, nbe3).
    As in the reference: the dimension-2 and dimension-3 derivatives and the
    2-3 cross derivatives are built along dimension 1 first (:889-1038, with
    the mixed derivatives involving dimension 1 as their Hermite data), the
    values along dimension 1 (:1058), then SplineBuilder2D over dimensions 2
    and 3 of the result with the built derivatives (:1060-1090)."""
    d = derivs or {}
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    nb = vals.shape[0]

    def b1(name, lo, hi):
        a = d.get(name)
        return None if a is None else _build_axis(s1, a, 1, d.get(lo), d.get(hi))

    d2min = b1("derivs_min2", "mixed_derivs_min1_min2", "mixed_derivs_max1_min2") if s2.bc_lower == HERMITE else None
    d2max = b1("derivs_max2", "mixed_derivs_min1_max2", "mixed_derivs_max1_max2") if s2.bc_upper == HERMITE else None
    d3min = b1("derivs_min3", "mixed_derivs_min1_min3", "mixed_derivs_max1_min3") if s3.bc_lower == HERMITE else None
    d3max = b1("derivs_max3", "mixed_derivs_min1_max3", "mixed_derivs_max1_max3") if s3.bc_upper == HERMITE else None
    m23 = {}
    for a2, bc2 in (("min", s2.bc_lower), ("max", s2.bc_upper)):
        for a3, bc3 in (("min", s3.bc_lower), ("max", s3.bc_upper)):
            if bc2 == HERMITE and bc3 == HERMITE:
                m23[a2, a3] = b1(f"mixed_derivs_{a2}2_{a3}3", f"mixed_derivs_min1_{a2}2_{a3}3",
                                 f"mixed_derivs_max1_{a2}2_{a3}3")
    spline1 = _build_axis(s1, vals, 1, d.get("derivs_min1"), d.get("derivs_max1"))

    def flat(a):
        return None if a is None else a.reshape((nb * s1.size,) + a.shape[2:])

    coef = spline2d_build(s2, s3, flat(spline1), flat(d2min), flat(d2max), flat(d3min), flat(d3max),
                          flat(m23.get(("min", "min"))), flat(m23.get(("max", "min"))),
                          flat(m23.get(("min", "max"))), flat(m23.get(("max", "max"))))
    return coef.reshape(nb, s1.size, s2.size, s3.size)


def _nd_args(splines, coef):
    nd = len(splines)
    hs = (C.c_void_p * nd)(*[s.h for s in splines])
    st = (C.c_long * nd)(*[coef.strides[k + 1] // 8 for k in range(nd)])
    return nd, hs, st


def splinend_eval(splines, coords, coef, orders=None, rules=None):
    """SplineEvaluatorND (spline_evaluator_nd.hpp:684-826): coords
    (nbatch, npts, N), coef (nbatch, size_1, ..., size_N) -> (nbatch, npts).
    rules: (lower, upper) per dimension, 0 = Null (periodic dimensions wrap)."""
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    coef = np.ascontiguousarray(coef, dtype=np.float64)
    nd, hs, st = _nd_args(splines, coef)
    nb, P = coords.shape[0], coords.shape[1]
    orders = (C.c_int * nd)(*(orders or [0] * nd))
    rules = (C.c_int * (2 * nd))(*(rules or [0] * (2 * nd)))
    out = np.empty((nb, P))
    if lib().or_splinend_eval(hs, nd, coords, out, P, coef, int(np.prod(coef.shape[1:])), st, nb, orders, rules):
        raise RuntimeError(_err())
    return out


def splinend_integrate(splines, coef):
    """SplineEvaluatorND::integrate (spline_evaluator_nd.hpp:597-632):
    coef (nbatch, size_1, ..., size_N) -> (nbatch,)."""
    coef = np.ascontiguousarray(coef, dtype=np.float64)
    nd, hs, st = _nd_args(splines, coef)
    out = np.empty(coef.shape[0])
    if lib().or_splinend_integrate(hs, nd, out, coef, int(np.prod(coef.shape[1:])), st, coef.shape[0]):
        raise RuntimeError(_err())
    return out
