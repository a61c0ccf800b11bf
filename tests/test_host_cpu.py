"""Host-side checks of the product library (no GPU compute): the C ABI exports
every entry point of include/ddcx.h, and the FP64 host setup (knots,
interpolation points, integrals, offsets, matrix shapes, quadrature weights,
validation errors) agrees with the oracle.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle as O
from cases import ALL_CASES, TILED_CASES, Case, breaks_for

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "ddcx.h")).read()
    return sorted(set(re.findall(r"\b(ddcx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_header_symbol():
    import ddc_amd

    lib = C.CDLL(ddc_amd.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s


def test_no_oracle_in_product():
    """The product library neither links to nor names the oracle."""
    import ddc_amd

    blob = open(ddc_amd.LIB_PATH, "rb").read()
    assert b"liboracle" not in blob and b"or_spline" not in blob


def test_version_string():
    import ddc_amd

    assert b"gfx950" in ddc_amd.lib().ddcx_version()


@pytest.mark.parametrize("case", ALL_CASES + TILED_CASES, ids=lambda c: c.id)
def test_host_setup_matches_oracle(case):
    orc = case.oracle()
    bs, builder = case.product()
    assert np.array_equal(bs.knots(), orc.knots())
    assert np.array_equal(builder.interpolation_points(), orc.points())
    import ddc_amd as D

    assert np.array_equal(D.integrals(bs), orc.integrals())
    assert builder.offset() == orc.offset
    assert builder.info.size == orc.size and builder.info.nbasis == orc.nbasis
    kinds = {0: "dense", 1: "band", 2: "pds_band", 3: "pds_tridiag"}
    blocked = {0: None, 1: "2x2", 2: "3x3"}[builder.info.blocked]
    assert (blocked or kinds[builder.info.matrix_kind]) == orc.matrix_type
    q1 = builder.quadrature_coefficients()
    q2 = orc.quadrature_coefficients()
    for a, b in zip(q1, q2):
        assert np.array_equal(a, b)


def test_errors_mirror_reference():
    import ddc_amd as D

    # check_valid_grid (spline_builder.hpp:1273-1317)
    bs = D.UniformBSplines(2, 0.0, 1.0, 5, periodic=True)
    for pts in ([0.1, 0.3, 0.5, 0.7], [0.1, 0.3, 0.5, 0.7, 0.9, 0.95], [0.1, 0.11, 0.12, 0.13, 0.14]):
        with pytest.raises(D.DDCXError) as e:
            D.SplineBuilder(bs, points=pts)
        assert e.value.status == 2
    # unsorted break points (non_uniform_point_sampling.hpp:88-90)
    with pytest.raises(D.DDCXError):
        D.NonUniformBSplines(3, [0.0, 0.5, 0.4, 1.0])
    # closures must both be periodic or both not
    with pytest.raises(D.DDCXError):
        D.SplineBuilder(bs, D.SplineBuilderClosure.PERIODIC, D.SplineBuilderClosure.GREVILLE)
    # degree outside the compiled range
    with pytest.raises(D.DDCXError) as e:
        D.UniformBSplines(9, 0.0, 1.0, 10)
    assert e.value.status == 5


def test_ginkgo_tag_routes_to_direct_solver():
    import ddc_amd as D

    bs = D.UniformBSplines(3, 0.0, 1.0, 32, periodic=True)
    b1 = D.SplineBuilder(bs, solver=D.SplineSolver.GINKGO)
    b2 = D.SplineBuilder(bs, solver=D.SplineSolver.LAPACK)
    assert b1.info.matrix_kind == b2.info.matrix_kind
    assert np.array_equal(b1.quadrature_coefficients()[1], b2.quadrature_coefficients()[1])


def test_batched_layout_helper():
    import torch

    import ddc_amd as D

    t = torch.empty((5, 1024, 7), dtype=torch.float64)
    L, nb = D.batched_layout(t, 1)
    assert (L.row_stride, L.inner_count, L.inner_stride, L.outer_stride, nb) == (7, 7, 1, 1024 * 7, 35)
    L, nb = D.batched_layout(t.permute(1, 0, 2), 0)
    assert nb == 35 and L.row_stride == 7


CPP_FACADE_PROGRAM = r"""
#include <cmath>
#include <cstdio>
#include "ddcx.hpp"
int main()
{
    std::vector<double> br{0.0, 0.05, 0.2, 0.3, 0.45, 0.5, 0.7, 0.8, 0.95, 1.0};
    ddcx::NonUniformBSplines bs(3, br, false);
    ddcx::SplineBuilder b(bs, ddcx::SplineBuilderClosure::GREVILLE, ddcx::SplineBuilderClosure::GREVILLE);
    auto pts = b.interpolation_points();
    auto q = b.quadrature_coefficients();
    double qs = 0.0;
    for (double v : q[1]) qs += v;
    std::printf("%d %d %.17g %.17g %.17g\n", b.npoints(), bs.size(), pts[1], pts[b.npoints() - 2], qs);
    ddcx::UniformBSplines ub(3, 0.0, 2.0, 16, true);
    ddcx::SplineBuilder pb(ub, ddcx::SplineBuilderClosure::PERIODIC, ddcx::SplineBuilderClosure::PERIODIC);
    std::printf("%d %d %d\n", pb.npoints(), pb.offset(), pb.info().kernel);
    try {
        std::vector<double> bad{0.0, 0.5, 0.4, 1.0};
        ddcx::NonUniformBSplines nb(3, bad, false);
        std::printf("no error\n");
    } catch (std::runtime_error const& e) {
        std::printf("error: %s\n", e.what());
    }
    int64_t nbatch = 0;
    ddcx_layout l = ddcx::batched_layout({5, 7, 11}, 1, &nbatch);
    std::printf("%lld %lld %lld %lld %lld\n", (long long)l.row_stride, (long long)l.inner_count,
                (long long)l.inner_stride, (long long)l.outer_stride, (long long)nbatch);
    for (int ax = 1; ax <= 3; ++ax) {
        ddcx_layout m = ddcx::batched_layout({2, 5, 6, 7}, ax, &nbatch);
        std::printf("%lld %lld %lld %lld %lld\n", (long long)m.row_stride, (long long)m.inner_count,
                    (long long)m.inner_stride, (long long)m.outer_stride, (long long)nbatch);
    }
    ddcx::SplineEvaluator e1(bs, ddcx::Extrapolation::Null, ddcx::Extrapolation::Null);
    ddcx::SplineEvaluator e2(ub, ddcx::Extrapolation::Periodic, ddcx::Extrapolation::Periodic);
    ddcx::SplineEvaluator3D e3(e1, e2, e1);
    try {
        e3.deriv({4, 0, 0}, nullptr, nullptr, 0, nullptr, 0, {1, 1, 1}, 0);
        std::printf("no error\n");
    } catch (std::runtime_error const& e) {
        std::printf("error: %s\n", e.what());
    }
    ddcx::SplineBuilder3D b3(b, pb, b, bs.size(), ub.size(), bs.size());
    std::printf("%lld %lld %d\n", (long long)b3.work1_size(2), (long long)b3.work2_size(2), e3.dimension());
    // a permuted (strided) array: extents (6, 5, 7), strides (7, 42, 1)
    ddcx_layout ps = ddcx::batched_layout({6, 5, 7}, {7, 42, 1}, 1, &nbatch);
    std::printf("%lld %lld %lld %lld %lld\n", (long long)ps.row_stride, (long long)ps.inner_count,
                (long long)ps.inner_stride, (long long)ps.outer_stride, (long long)nbatch);
    try {
        ddcx::SplineEvaluator bad(bs, ddcx::ConstantExtrapolationRule(2.0), ddcx::NullExtrapolationRule());
        std::printf("no error\n");
    } catch (std::runtime_error const& e) {
        std::printf("error: %s\n", e.what());
    }
    ddcx::SplineEvaluator ok(bs, ddcx::ConstantExtrapolationRule(0.25, 0.0, 1.0), ddcx::ConstantExtrapolationRule(0.75));
    ddcx::SplineBuilder2D b2d(b, pb, bs.size());
    std::printf("%lld\n", (long long)b2d.workspace_size(3));
    return 0;
}
"""


def test_cpp_facade(tmp_path):
    """include/ddcx.hpp (the C++20 facade with the reference's class names)
    compiles against the C ABI and answers host-side queries like the Python
    mirror; errors surface as std::runtime_error."""
    import shutil
    import subprocess

    import ddc_amd as D

    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    src = tmp_path / "facade.cpp"
    src.write_text(CPP_FACADE_PROGRAM)
    exe = tmp_path / "facade"
    libdir = os.path.dirname(D.LIB_PATH)
    subprocess.run([cxx, "-std=c++20", "-O1", str(src), "-I", os.path.join(ROOT, "include"), "-L", libdir,
                    "-lddcx", f"-Wl,-rpath,{libdir}", "-Wl,--allow-shlib-undefined", "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines()
    br = np.array([0.0, 0.05, 0.2, 0.3, 0.45, 0.5, 0.7, 0.8, 0.95, 1.0])
    bs = D.NonUniformBSplines(3, br)
    b = D.SplineBuilder(bs, D.SplineBuilderClosure.GREVILLE, D.SplineBuilderClosure.GREVILLE)
    pts = b.interpolation_points()
    ql, qv, qr = b.quadrature_coefficients()
    f = out[0].split()
    assert int(f[0]) == b.npoints() and int(f[1]) == bs.size()
    assert float(f[2]) == pts[1] and float(f[3]) == pts[-2]
    assert abs(float(f[4]) - qv.sum()) <= 1e-15
    ub = D.UniformBSplines(3, 0.0, 2.0, 16, periodic=True)
    pb = D.SplineBuilder(ub)
    assert out[1].split() == [str(pb.npoints()), str(pb.offset()), str(pb.info.kernel)]
    assert out[2].startswith("error:")
    assert out[3].split() == ["11", "11", "1", "77", "55"]  # same as splines.batched_layout
    import torch
    t = torch.empty((2, 5, 6, 7), dtype=torch.float64)
    for ax in (1, 2, 3):
        lay, nb = D.batched_layout(t, ax)
        assert out[3 + ax].split() == [str(v) for v in (lay.row_stride, lay.inner_count, lay.inner_stride,
                                                       lay.outer_stride, nb)]
    assert out[7].startswith("error:") and "order" in out[7]
    n1, s1 = b.npoints(), bs.size()
    assert out[8].split() == [str(2 * s1 * pb.npoints() * n1), str(2 * s1 * ub.size() * n1), "3"]
    tp = torch.empty((5, 6, 7), dtype=torch.float64).permute(1, 0, 2)
    lay, nb = D.batched_layout(tp, 1)
    assert out[9].split() == [str(v) for v in (lay.row_stride, lay.inner_count, lay.inner_stride,
                                               lay.outer_stride, nb)]
    assert out[10].startswith("error:") and "inside the domain" in out[10]
    assert out[11] == str(3 * s1 * pb.npoints())


def test_knots_as_interpolation_points():
    """KnotsAsInterpolationPoints: break points, ncells + 1 of them (ncells
    when periodic); uniform ones as UniformPointSampling::init samples them;
    Greville closures refused."""
    import ddc_amd as D

    C = D.SplineBuilderClosure
    u = D.UniformBSplines(3, -1.0, 2.0, 10)
    p = D.knots_as_interpolation_points(u, C.HERMITE, C.HERMITE)
    assert p.size == 11 and p[0] == -1.0
    assert np.array_equal(p, -1.0 + np.arange(11) * (3.0 / 10))
    up = D.UniformBSplines(3, 0.0, 1.0, 10, periodic=True)
    assert D.knots_as_interpolation_points(up, C.PERIODIC, C.PERIODIC).size == 10
    br = np.array([0.0, 0.1, 0.35, 0.5, 0.9, 1.0])
    nu = D.NonUniformBSplines(2, br)
    assert np.array_equal(D.knots_as_interpolation_points(nu, C.HOMOGENEOUS_HERMITE, C.HERMITE), br)
    with pytest.raises(ValueError):
        D.knots_as_interpolation_points(nu, C.GREVILLE, C.HERMITE)


def test_linear_problem_host_side():
    """ddcx_problem_*: the maker's storage choices, set/get before and after
    setup, setup's factorisation, and the argument errors (no device work)."""
    import ddc_amd.linear_problem as LP
    from ddc_amd import DDCXError
    from matrices import nonsym

    assert LP.make_new_band(10, 3, 3, False).main_kind() == "dense"  # 2 kl + ku + 1 >= n
    assert LP.make_new_band(20, 3, 3, False).main_kind() == "band"
    assert LP.make_new_band(20, 1, 1, True).main_kind() == "pds_tridiag"
    assert LP.make_new_band(20, 2, 2, True).main_kind() == "pds_band"
    assert LP.make_new_band(20, 0, 4, True).main_kind() == "band"  # kl != ku: not symmetric storage
    assert LP.make_new_band(10, 0, 12, True).main_kind() == "dense"  # the maker's fallback takes ku >= n
    with pytest.raises(DDCXError):
        LP.make_new_band(10, 12, 12, True)  # symmetric band storage wider than the matrix
    p = LP.make_new_block_matrix_with_band_main_block(20, 2, 2, False, 3, 2)
    assert p.blocked() == 2 and p.required_number_of_rhs_rows() == 22
    assert LP.make_new_periodic_band_matrix(10, 4, 4, False).blocked() == 0  # dense is smaller
    A = nonsym(20, 2)
    p = LP.make_new_block_matrix_with_band_main_block(20, 2, 2, False, 3)
    for i in range(20):
        for j in range(20):
            if A[i, j]:
                p.set_element(i, j, A[i, j])
    got = np.array([[p.get_element(i, j) for j in range(20)] for i in range(20)])
    assert np.array_equal(got, A)
    p.setup_solver()
    got = np.array([[p.get_element(i, j) for j in range(20)] for i in range(20)])
    assert np.array_equal(got, A)
    with pytest.raises(DDCXError):
        p.set_element(0, 0, 1.0)
    with pytest.raises(DDCXError):
        p.get_element(20, 0)
    with pytest.raises(DDCXError):
        LP.make_new_block_matrix_with_band_main_block(5, 1, 1, False, 3, 2)  # no main block left
    q = LP.make_new_dense(3)
    with pytest.raises(DDCXError, match="SINGULAR"):
        q.setup_solver()  # all zeros: dgetrf info > 0
