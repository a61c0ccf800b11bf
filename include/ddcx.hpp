// ddcx.hpp -- C++20 facade over the libddcx C ABI (ddcx.h) with the names and
// call shapes of DDC's spline classes: ddcx::UniformBSplines /
// NonUniformBSplines (bsplines_uniform.hpp:48-370,
// bsplines_non_uniform.hpp:44-392), ddcx::SplineBuilder
// (spline_builder.hpp:57-617) and ddcx::SplineEvaluator
// (spline_evaluator.hpp:33-606).  Errors are rethrown as std::runtime_error,
// as the reference throws.  Device arrays are described by ddcx::Span, a
// pointer plus the strides of a layout_right array with one interpolation
// dimension and up to two batch dimensions (what a ddc::ChunkSpan's mdspan
// provides through stride()).
#pragma once

#include <cstdint>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

#include "ddcx.h"

namespace ddcx {

inline void check(ddcx_status s)
{
    if (s != DDCX_OK) throw std::runtime_error(ddcx_last_error());
}

enum class SplineBuilderClosure { PERIODIC = DDCX_PERIODIC, HERMITE = DDCX_HERMITE,
                                  HOMOGENEOUS_HERMITE = DDCX_HOMOGENEOUS_HERMITE, GREVILLE = DDCX_GREVILLE };
enum class SplineSolver { GINKGO = DDCX_SOLVER_GINKGO, LAPACK = DDCX_SOLVER_LAPACK };
enum class Precision { FP64 = DDCX_FP64, FP32_REFINE = DDCX_FP32_REFINE };
enum class Extrapolation { Null = DDCX_EXTRAP_NULL, Constant = DDCX_EXTRAP_CONSTANT, Periodic = DDCX_EXTRAP_PERIODIC };

// A batched device array: element (row r, batch line l) at
// data[r*row_stride + (l % inner_count)*inner_stride + (l / inner_count)*outer_stride].
template <class T>
struct Span {
    T* data;
    ddcx_layout layout;
    int64_t nbatch;
};

// Layout of a layout_right array whose extents are `ext` (outermost first)
// and whose interpolation dimension is `axis` (at most two other dimensions).
inline ddcx_layout batched_layout(std::vector<int64_t> const& ext, int axis, int64_t* nbatch = nullptr)
{
    int const nd = int(ext.size());
    std::vector<int64_t> st(nd, 1);
    for (int d = nd - 2; d >= 0; --d) st[d] = st[d + 1] * ext[d + 1];
    std::vector<int> others;
    for (int d = 0; d < nd; ++d)
        if (d != axis) others.push_back(d);
    if (others.size() > 2) throw std::runtime_error("at most two batch dimensions");
    ddcx_layout l{st[axis], 1, 0, 0};
    int64_t nb = 1;
    if (others.size() == 1) {
        l.inner_count = ext[others[0]];
        l.inner_stride = st[others[0]];
        nb = ext[others[0]];
    } else if (others.size() == 2) {
        l.inner_count = ext[others[1]];
        l.inner_stride = st[others[1]];
        l.outer_stride = st[others[0]];
        nb = ext[others[0]] * ext[others[1]];
    }
    if (nbatch) *nbatch = nb;
    return l;
}

class BSplines {
public:
    int degree() const { return desc_.degree; }
    bool is_periodic() const { return desc_.periodic != 0; }
    bool is_uniform() const { return desc_.uniform != 0; }
    int ncells() const { return desc_.ncells; }
    int nbasis() const { return desc_.ncells + (desc_.periodic ? 0 : desc_.degree); }
    int size() const { return desc_.ncells + desc_.degree; }
    std::vector<double> integrals() const
    {
        std::vector<double> v(size());
        check(ddcx_bsplines_integrals(h_.get(), v.data()));
        return v;
    }
    ddcx_bsplines* handle() const { return h_.get(); }

protected:
    BSplines(ddcx_bsplines_desc const& d, std::vector<double> breaks) : desc_(d), breaks_(std::move(breaks))
    {
        desc_.breaks = breaks_.empty() ? nullptr : breaks_.data();
        ddcx_bsplines* h = nullptr;
        check(ddcx_bsplines_create(&desc_, &h));
        h_ = std::shared_ptr<ddcx_bsplines>(h, Del{});
    }

private:
    struct Del {
        void operator()(ddcx_bsplines* h) const { ddcx_bsplines_destroy(h); }
    };
    ddcx_bsplines_desc desc_;
    std::vector<double> breaks_;
    std::shared_ptr<ddcx_bsplines> h_;
};

// UniformBSplines<CDim, D>::Impl::init(rmin, rmax, ncells), bsplines_uniform.hpp:135-148
class UniformBSplines : public BSplines {
public:
    UniformBSplines(int degree, double rmin, double rmax, int ncells, bool periodic)
        : BSplines(ddcx_bsplines_desc{degree, 1, periodic ? 1 : 0, ncells, rmin, rmax, nullptr}, {})
    {
    }
};

// NonUniformBSplines<CDim, D>::Impl::init(breaks), bsplines_non_uniform.hpp:414-457
class NonUniformBSplines : public BSplines {
public:
    NonUniformBSplines(int degree, std::vector<double> breaks, bool periodic)
        : BSplines(ddcx_bsplines_desc{degree, 0, periodic ? 1 : 0, int(breaks.size()) - 1, breaks.front(),
                                      breaks.back(), nullptr},
                   breaks)
    {
    }
};

// SplineBuilder<Exec, Mem, BSplines, DDimI, BCl, BCu, Solver>
class SplineBuilder {
public:
    SplineBuilder(BSplines const& bs, SplineBuilderClosure lower, SplineBuilderClosure upper,
                  SplineSolver solver = SplineSolver::LAPACK, Precision precision = Precision::FP64,
                  std::vector<double> const* points = nullptr)
    {
        ddcx_builder_desc d{ddcx_closure(lower), ddcx_closure(upper), ddcx_solver(solver),
                            ddcx_precision(precision), points ? points->data() : nullptr,
                            points ? int(points->size()) : 0, 0};
        ddcx_builder* h = nullptr;
        check(ddcx_builder_create(bs.handle(), &d, &h));
        h_.reset(h);
        check(ddcx_builder_get_info(h, &info_));
    }
    int npoints() const { return info_.npoints; }
    int offset() const { return info_.offset; }
    ddcx_builder_info const& info() const { return info_; }
    std::vector<double> interpolation_points() const
    {
        std::vector<double> p(info_.npoints);
        check(ddcx_builder_points(h_.get(), p.data()));
        return p;
    }
    // operator()(spline, vals, derivs_xmin, derivs_xmax), spline_builder.hpp:538-557
    void operator()(Span<double> spline, Span<double const> vals,
                    std::optional<Span<double const>> derivs_xmin = std::nullopt,
                    std::optional<Span<double const>> derivs_xmax = std::nullopt, void* stream = nullptr) const
    {
        ddcx_layout dl{1, 1, 0, 0};
        if (derivs_xmin) dl = derivs_xmin->layout;
        if (derivs_xmax) dl = derivs_xmax->layout;
        check(ddcx_builder_build(h_.get(), spline.data, spline.layout, vals.data, vals.layout,
                                 derivs_xmin ? derivs_xmin->data : nullptr, derivs_xmax ? derivs_xmax->data : nullptr,
                                 dl, vals.nbatch, stream));
    }
    // adjoint of operator(): gvals (and Hermite derivative gradients) from gspline
    void build_vjp(Span<double const> gspline, Span<double> gvals, std::optional<Span<double>> gderivs_xmin = {},
                   std::optional<Span<double>> gderivs_xmax = {}, void* stream = nullptr) const
    {
        ddcx_layout dl{1, 1, 0, 0};
        if (gderivs_xmin) dl = gderivs_xmin->layout;
        if (gderivs_xmax) dl = gderivs_xmax->layout;
        check(ddcx_builder_build_vjp(h_.get(), gspline.data, gspline.layout, gvals.data, gvals.layout,
                                     gderivs_xmin ? gderivs_xmin->data : nullptr,
                                     gderivs_xmax ? gderivs_xmax->data : nullptr, dl, gvals.nbatch, stream));
    }
    // quadrature_coefficients(), spline_builder.hpp:1111-1234: {derivs_xmin, vals, derivs_xmax}
    std::vector<std::vector<double>> quadrature_coefficients() const
    {
        std::vector<double> ql(info_.nbv_lower), qv(info_.nbasis - info_.nbv_lower - info_.nbv_upper),
                qr(info_.nbv_upper);
        check(ddcx_builder_quadrature(h_.get(), ql.data(), qv.data(), qr.data()));
        return {ql, qv, qr};
    }

private:
    struct Del {
        void operator()(ddcx_builder* h) const { ddcx_builder_destroy(h); }
    };
    std::unique_ptr<ddcx_builder, Del> h_;
    ddcx_builder_info info_{};
};

// SplineEvaluator<Exec, Mem, BSplines, EvalDDim, LowerRule, UpperRule>
class SplineEvaluator {
public:
    SplineEvaluator(BSplines const& bs, Extrapolation lower, Extrapolation upper)
    {
        ddcx_evaluator* h = nullptr;
        check(ddcx_evaluator_create(bs.handle(), ddcx_extrap(lower), ddcx_extrap(upper), &h));
        h_.reset(h);
    }
    // operator()(spline_eval, coords_eval, spline_coef), spline_evaluator.hpp:263
    void operator()(Span<double> out, Span<double const> coords, Span<double const> coef, int64_t npts,
                    void* stream = nullptr) const
    {
        check(ddcx_evaluator_eval(h_.get(), out.data, coords.data, coords.layout, npts, coef.data, coef.layout,
                                  coef.nbatch, stream));
    }
    // deriv(order, spline_eval, coords_eval, spline_coef), spline_evaluator.hpp:395
    void deriv(int order, Span<double> out, Span<double const> coords, Span<double const> coef, int64_t npts,
               void* stream = nullptr) const
    {
        check(ddcx_evaluator_deriv(h_.get(), order, out.data, coords.data, coords.layout, npts, coef.data,
                                   coef.layout, coef.nbatch, stream));
    }
    // integrate(integrals, spline_coef), spline_evaluator.hpp:494
    void integrate(double* out, int64_t out_stride, Span<double const> coef, void* stream = nullptr) const
    {
        check(ddcx_evaluator_integrate(h_.get(), out, out_stride, coef.data, coef.layout, coef.nbatch, stream));
    }
    void eval_vjp(Span<double const> coords, Span<double const> w, Span<double> gx, Span<double const> coef,
                  Span<double> gcoef, int64_t npts, void* stream = nullptr) const
    {
        (void)w.layout;
        check(ddcx_evaluator_eval_vjp(h_.get(), coords.data, w.data, gx.data, coords.layout, npts, coef.data,
                                      gcoef.data, coef.layout, coef.nbatch, stream));
    }
    ddcx_evaluator* handle() const { return h_.get(); }

private:
    struct Del {
        void operator()(ddcx_evaluator* h) const { ddcx_evaluator_destroy(h); }
    };
    std::unique_ptr<ddcx_evaluator, Del> h_;
};

} // namespace ddcx
