// ddcx.hpp -- C++20 facade over the libddcx C ABI (ddcx.h) with the names and
// call shapes of DDC's spline classes: ddcx::UniformBSplines /
// NonUniformBSplines (bsplines_uniform.hpp:48-370,
// bsplines_non_uniform.hpp:44-392), ddcx::SplineBuilder
// (spline_builder.hpp:57-617), ddcx::SplineEvaluator
// (spline_evaluator.hpp:33-606) with the extrapolation rules
// (null_extrapolation_rule.hpp, constant_extrapolation_rule.hpp,
// periodic_extrapolation_rule.hpp), ddcx::SplineBuilder2D /
// SplineEvaluator2D (spline_builder_2d.hpp:27-651,
// spline_evaluator_2d.hpp:37-826), SplineBuilder3D / SplineEvaluator3D /
// SplineEvaluatorND.  Errors are rethrown as std::runtime_error,
// as the reference throws.  Device arrays are described by ddcx::Span, a
// pointer plus the strides of a layout_right array with one interpolation
// dimension and up to two batch dimensions (what a ddc::ChunkSpan's mdspan
// provides through stride()).
#pragma once

#include <cmath>
#include <cstdint>
#include <algorithm>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
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
// and whose interpolation dimension is `axis`.  The other dimensions are batch
// dimensions; neighbouring ones on the same side of `axis` merge (they are
// contiguous in a layout_right array), so at most two remain.
inline ddcx_layout batched_layout(std::vector<int64_t> const& ext, int axis, int64_t* nbatch = nullptr)
{
    int const nd = int(ext.size());
    std::vector<int64_t> st(nd, 1);
    for (int d = nd - 2; d >= 0; --d) st[d] = st[d + 1] * ext[d + 1];
    std::vector<std::pair<int64_t, int64_t>> groups; // (count, stride), outermost first
    auto side = [&](int d0, int d1) {
        if (d0 >= d1) return;
        int64_t n = 1;
        for (int d = d0; d < d1; ++d) n *= ext[d];
        groups.emplace_back(n, st[d1 - 1]);
    };
    side(0, axis);
    side(axis + 1, nd);
    ddcx_layout l{st[axis], 1, 0, 0};
    int64_t nb = 1;
    if (groups.size() == 1) {
        l.inner_count = groups[0].first;
        l.inner_stride = groups[0].second;
        nb = groups[0].first;
    } else if (groups.size() == 2) {
        l.inner_count = groups[1].first;
        l.inner_stride = groups[1].second;
        l.outer_stride = groups[0].second;
        nb = groups[0].first * groups[1].first;
    }
    if (nbatch) *nbatch = nb;
    return l;
}

// Layout of a strided array (what a ChunkSpan's mdspan gives through
// extent(d) and stride(d), outermost first) whose interpolation dimension is
// `axis`.  Neighbouring batch dimensions on the same side of `axis` merge
// when they are contiguous with each other (stride[d] == ext[d+1] *
// stride[d+1]); at most two batch dimensions may remain.
inline ddcx_layout batched_layout(std::vector<int64_t> const& ext, std::vector<int64_t> const& stride, int axis,
                                  int64_t* nbatch = nullptr)
{
    int const nd = int(ext.size());
    if (int(stride.size()) != nd || axis < 0 || axis >= nd) throw std::runtime_error("ddcx: bad extents or axis");
    std::vector<std::pair<int64_t, int64_t>> groups; // (count, stride), outermost first
    auto side = [&](int d0, int d1) {
        bool open = false;
        for (int d = d0; d < d1; ++d) {
            if (open && groups.back().second == ext[d] * stride[d]) {
                groups.back() = {groups.back().first * ext[d], stride[d]};
            } else {
                groups.emplace_back(ext[d], stride[d]);
                open = true;
            }
        }
    };
    side(0, axis);
    side(axis + 1, nd);
    if (groups.size() > 2) throw std::runtime_error("ddcx: at most two batch dimensions (after merging)");
    ddcx_layout l{stride[axis], 1, 0, 0};
    int64_t nb = 1;
    if (groups.size() == 1) {
        l.inner_count = groups[0].first;
        l.inner_stride = groups[0].second;
        nb = groups[0].first;
    } else if (groups.size() == 2) {
        l.inner_count = groups[1].first;
        l.inner_stride = groups[1].second;
        l.outer_stride = groups[0].second;
        nb = groups[0].first * groups[1].first;
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
    ddcx_builder* handle() const { return h_.get(); }

private:
    struct Del {
        void operator()(ddcx_builder* h) const { ddcx_builder_destroy(h); }
    };
    std::unique_ptr<ddcx_builder, Del> h_;
    ddcx_builder_info info_{};
};

// Extrapolation rule objects, passed to the SplineEvaluator constructors as
// the reference passes its rule objects (spline_evaluator.hpp:140-153).
struct ExtrapolationRule {
    ddcx_extrap_rule rule;
};
// NullExtrapolationRule (null_extrapolation_rule.hpp:14-26): 0 outside the domain
inline ExtrapolationRule NullExtrapolationRule() { return {{DDCX_EXTRAP_NULL, NAN, NAN, NAN}}; }
// PeriodicExtrapolationRule<DimI> (periodic_extrapolation_rule.hpp:12-24)
inline ExtrapolationRule PeriodicExtrapolationRule() { return {{DDCX_EXTRAP_PERIODIC, NAN, NAN, NAN}}; }
// ConstantExtrapolationRule<DimI>(eval_pos) (constant_extrapolation_rule.hpp:39-40)
inline ExtrapolationRule ConstantExtrapolationRule(double eval_pos) { return {{DDCX_EXTRAP_CONSTANT, eval_pos, NAN, NAN}}; }
// ConstantExtrapolationRule<DimI, DimNI>(eval_pos, eval_pos_not_interest_min,
// eval_pos_not_interest_max) (constant_extrapolation_rule.hpp:99-110), for
// the evaluators of a SplineEvaluator2D
inline ExtrapolationRule ConstantExtrapolationRule(double eval_pos, double eval_pos_not_interest_min,
                                                   double eval_pos_not_interest_max)
{
    return {{DDCX_EXTRAP_CONSTANT, eval_pos, eval_pos_not_interest_min, eval_pos_not_interest_max}};
}

// SplineEvaluator<Exec, Mem, BSplines, EvalDDim, LowerRule, UpperRule>
class SplineEvaluator {
public:
    // SplineEvaluator(lower_extrap_rule, upper_extrap_rule), spline_evaluator.hpp:140-153
    SplineEvaluator(BSplines const& bs, ExtrapolationRule const& lower, ExtrapolationRule const& upper)
    {
        ddcx_evaluator* h = nullptr;
        check(ddcx_evaluator_create_with_rules(bs.handle(), &lower.rule, &upper.rule, &h));
        h_.reset(h);
    }
    // rules by kind; a Constant rule evaluates at the bound it guards
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
    // operator()(spline_eval, spline_coef) / deriv(order, spline_eval,
    // spline_coef) (spline_evaluator.hpp:296-342, :432-478): at the
    // evaluation domain's own coordinates `points` (npts of them, device
    // memory), the same for every line; order 0: operator()
    void eval_at(int order, Span<double> out, const double* points, int64_t npts, Span<double const> coef,
                 void* stream = nullptr) const
    {
        if (out.nbatch != coef.nbatch) throw std::runtime_error("ddcx: batch sizes differ");
        check(ddcx_evaluator_eval_at(h_.get(), order, out.data, out.layout, points, npts, coef.data, coef.layout,
                                     out.nbatch, stream));
    }
    // integrate(integrals, spline_coef), spline_evaluator.hpp:494
    void integrate(double* out, int64_t out_stride, Span<double const> coef, void* stream = nullptr) const
    {
        check(ddcx_evaluator_integrate(h_.get(), out, out_stride, coef.data, coef.layout, coef.nbatch, stream));
    }
    // adjoint of operator(): gcoef = sum_p w_p dB(x_p)/dcoef, gx_p = w_p f'(x_p);
    // w and gx must have the layout of coords, gcoef that of coef
    void eval_vjp(Span<double const> coords, Span<double const> w, Span<double> gx, Span<double const> coef,
                  Span<double> gcoef, int64_t npts, void* stream = nullptr) const
    {
        if (!same(w.layout, coords.layout) || !same(gx.layout, coords.layout))
            throw std::runtime_error("ddcx: eval_vjp needs w and gx in the layout of coords");
        if (!same(gcoef.layout, coef.layout))
            throw std::runtime_error("ddcx: eval_vjp needs gcoef in the layout of coef");
        check(ddcx_evaluator_eval_vjp(h_.get(), coords.data, w.data, gx.data, coords.layout, npts, coef.data,
                                      gcoef.data, coef.layout, coef.nbatch, stream));
    }
    ddcx_evaluator* handle() const { return h_.get(); }
    // the device data for the device-callable single-point evaluation of
    // ddcx_device.hpp (operator()(coord, coef_1d), deriv(order, coord, coef_1d))
    ddcx_device_evaluator device_view() const
    {
        ddcx_device_evaluator d{};
        check(ddcx_evaluator_device_view(h_.get(), &d));
        return d;
    }

private:
    static bool same(ddcx_layout const& a, ddcx_layout const& b)
    {
        return a.row_stride == b.row_stride && a.inner_count == b.inner_count && a.inner_stride == b.inner_stride &&
               a.outer_stride == b.outer_stride;
    }
    struct Del {
        void operator()(ddcx_evaluator* h) const { ddcx_evaluator_destroy(h); }
    };
    std::unique_ptr<ddcx_evaluator, Del> h_;
};

// builder(spline, vals, derivs...) then evaluator(out, coords, spline) on the
// same lines in one pass (ddcx_build_eval).  spline.data may be null: the
// coefficients then stay on chip (fused kernel only).
inline void build_eval(SplineBuilder const& b, SplineEvaluator const& e, Span<double> spline,
                       Span<double const> vals, Span<double> out, Span<double const> coords, int64_t npts,
                       std::optional<Span<double const>> derivs_xmin = std::nullopt,
                       std::optional<Span<double const>> derivs_xmax = std::nullopt, void* stream = nullptr)
{
    ddcx_layout dl{1, 1, 0, 0};
    if (derivs_xmin) dl = derivs_xmin->layout;
    if (derivs_xmax) dl = derivs_xmax->layout;
    check(ddcx_build_eval(b.handle(), e.handle(), spline.data, spline.layout, vals.data, vals.layout,
                          derivs_xmin ? derivs_xmin->data : nullptr, derivs_xmax ? derivs_xmax->data : nullptr, dl,
                          out.data, coords.data, coords.layout, npts, vals.nbatch, stream));
}

// One semi-Lagrangian advection step (examples/characteristics_advection.cpp:
// 237-274) in one fused kernel (ddcx_advect): out = the spline of vals at the
// feet x_i - shift[l]; shift holds one value per batch line (device memory).
// spline.data may be null (coefficients kept on chip).
inline void advect(SplineBuilder const& b, SplineEvaluator const& e, Span<double> out, Span<double const> vals,
                   const double* shift, Span<double> spline = {nullptr, {1, 1, 0, 0}, 0},
                   std::optional<Span<double const>> derivs_xmin = std::nullopt,
                   std::optional<Span<double const>> derivs_xmax = std::nullopt, void* stream = nullptr)
{
    ddcx_layout dl{1, 1, 0, 0};
    if (derivs_xmin) dl = derivs_xmin->layout;
    if (derivs_xmax) dl = derivs_xmax->layout;
    check(ddcx_advect(b.handle(), e.handle(), spline.data, spline.layout, vals.data, vals.layout,
                      derivs_xmin ? derivs_xmin->data : nullptr, derivs_xmax ? derivs_xmax->data : nullptr, dl, shift,
                      out.data, out.layout, vals.nbatch, stream));
}

// SplineBuilder2D<Exec, Mem, BSplines1, BSplines2, DDimI1, DDimI2, BCs...,
// Solver> (spline_builder_2d.hpp:27-651) over one 1D builder per dimension
// (not owned), on layout_right device arrays: vals (nbatch, n1, n2) ->
// spline (nbatch, size1, size2).  With Hermite closures: derivs_min1 /
// derivs_max1 (nbatch, nbe1, n2) along dimension 1 at each dimension-2 point,
// derivs_min2 / derivs_max2 (nbatch, n1, nbe2), and the corner cross
// derivatives mixed_* (nbatch, nbe1, nbe2).  The dimension-2 derivatives are
// built along dimension 1 first (with the corner derivatives as their
// Hermite data), then dimension 2 over the dimension-1 spline (:514-651).
// The intermediates live in a caller workspace of workspace_size(nbatch)
// doubles; nothing is allocated.
class SplineBuilder2D {
public:
    SplineBuilder2D(SplineBuilder const& b1, SplineBuilder const& b2, int64_t size1) : b1_(&b1), b2_(&b2), s1_(size1)
    {
    }
    int64_t nbe1() const { return std::max(b1_->info().nbv_lower, b1_->info().nbv_upper); }
    int64_t nbe2() const { return std::max(b2_->info().nbv_lower, b2_->info().nbv_upper); }
    int64_t workspace_size(int64_t nbatch) const
    {
        return nbatch * s1_ * (b2_->npoints() + b2_->info().nbv_lower + b2_->info().nbv_upper);
    }
    void operator()(double* spline, const double* vals, int64_t nbatch, double* workspace,
                    const double* derivs_min1 = nullptr, const double* derivs_max1 = nullptr,
                    const double* derivs_min2 = nullptr, const double* derivs_max2 = nullptr,
                    const double* mixed_min1_min2 = nullptr, const double* mixed_max1_min2 = nullptr,
                    const double* mixed_min1_max2 = nullptr, const double* mixed_max1_max2 = nullptr,
                    void* stream = nullptr) const
    {
        int64_t const n1 = b1_->npoints(), n2 = b2_->npoints(), s2 = b2_->info().size;
        int64_t const l2 = b2_->info().nbv_lower, u2 = b2_->info().nbv_upper, e1 = nbe1();
        double* spline1 = workspace;
        double* d1min = spline1 + nbatch * s1_ * n2;
        double* d1max = d1min + nbatch * s1_ * l2;
        auto dim1 = [&](double* out, const double* in, int64_t cols, const double* dmin, const double* dmax) {
            int64_t nb = 0;
            ddcx_layout const lo = batched_layout({nbatch, s1_, cols}, 1, &nb);
            ddcx_layout const li = batched_layout({nbatch, n1, cols}, 1);
            ddcx_layout const ld = batched_layout({nbatch, std::max<int64_t>(e1, 1), cols}, 1);
            check(ddcx_builder_build(b1_->handle(), out, lo, in, li, dmin, dmax, ld, nb, stream));
        };
        if (l2 > 0) dim1(d1min, derivs_min2, l2, mixed_min1_min2, mixed_max1_min2);
        dim1(spline1, vals, n2, derivs_min1, derivs_max1);
        if (u2 > 0) dim1(d1max, derivs_max2, u2, mixed_min1_max2, mixed_max1_max2);
        int64_t nb = 0;
        ddcx_layout const lo = batched_layout({nbatch, s1_, s2}, 2, &nb);
        ddcx_layout const li = batched_layout({nbatch, s1_, n2}, 2);
        ddcx_layout const ld = batched_layout({nbatch, s1_, std::max<int64_t>(nbe2(), 1)}, 2);
        check(ddcx_builder_build(b2_->handle(), spline, lo, spline1, li, l2 > 0 ? d1min : nullptr,
                                 u2 > 0 ? d1max : nullptr, ld, nb, stream));
    }
    // adjoint of operator(): gvals and, with Hermite closures, the gradients
    // of the derivative inputs (shaped as operator()'s; required where the
    // closures take derivatives), from gspline, through the same workspace:
    // the transposed dimension-2 solve gives the gradients of the dimension-1
    // splines, then the transposed dimension-1 solves of operator()'s three
    // dimension-1 builds give the rest
    void build_vjp(const double* gspline, double* gvals, int64_t nbatch, double* workspace,
                   double* gderivs_min1 = nullptr, double* gderivs_max1 = nullptr, double* gderivs_min2 = nullptr,
                   double* gderivs_max2 = nullptr, double* gmixed_min1_min2 = nullptr,
                   double* gmixed_max1_min2 = nullptr, double* gmixed_min1_max2 = nullptr,
                   double* gmixed_max1_max2 = nullptr, void* stream = nullptr) const
    {
        int64_t const n1 = b1_->npoints(), n2 = b2_->npoints(), s2 = b2_->info().size;
        int64_t const l2 = b2_->info().nbv_lower, u2 = b2_->info().nbv_upper, e1 = nbe1();
        double* gs1 = workspace;
        double* gd1min = gs1 + nbatch * s1_ * n2;
        double* gd1max = gd1min + nbatch * s1_ * l2;
        int64_t nb = 0;
        ddcx_layout const lg = batched_layout({nbatch, s1_, s2}, 2, &nb);
        ddcx_layout const lt = batched_layout({nbatch, s1_, n2}, 2);
        ddcx_layout const ld = batched_layout({nbatch, s1_, std::max<int64_t>(nbe2(), 1)}, 2);
        check(ddcx_builder_build_vjp(b2_->handle(), gspline, lg, gs1, lt, l2 > 0 ? gd1min : nullptr,
                                     u2 > 0 ? gd1max : nullptr, ld, nb, stream));
        auto dim1 = [&](const double* gin, double* gout, int64_t cols, double* gdmin, double* gdmax) {
            int64_t nb1 = 0;
            ddcx_layout const li = batched_layout({nbatch, s1_, cols}, 1, &nb1);
            ddcx_layout const lo = batched_layout({nbatch, n1, cols}, 1);
            ddcx_layout const lgd = batched_layout({nbatch, std::max<int64_t>(e1, 1), cols}, 1);
            check(ddcx_builder_build_vjp(b1_->handle(), gin, li, gout, lo, gdmin, gdmax, lgd, nb1, stream));
        };
        dim1(gs1, gvals, n2, gderivs_min1, gderivs_max1);
        if (l2 > 0) dim1(gd1min, gderivs_min2, l2, gmixed_min1_min2, gmixed_max1_min2);
        if (u2 > 0) dim1(gd1max, gderivs_max2, u2, gmixed_min1_max2, gmixed_max1_max2);
    }

private:
    SplineBuilder const* b1_;
    SplineBuilder const* b2_;
    int64_t s1_;
};

// SplineEvaluator2D<Exec, Mem, BSplines1, BSplines2, DDim1, DDim2, rules...>
// (spline_evaluator_2d.hpp:37-826) over one 1D evaluator per dimension (not
// owned; their rules are the 2D rules: lower/upper of dimension 1, then of
// dimension 2).  Points interleave (x1, x2): point p of batch l at
// coords[2*(l*npts + p)], out[l*npts + p]; coefficients of batch l start at
// coef + l*coef_batch_stride, element (i1, i2) at i1*st1 + i2*st2.
class SplineEvaluator2D {
public:
    SplineEvaluator2D(SplineEvaluator const& e1, SplineEvaluator const& e2) : e1_(&e1), e2_(&e2) {}
    // operator()(spline_eval, coords_eval, spline_coef), spline_evaluator_2d.hpp:370-420
    void operator()(double* out, const double* coords, int64_t npts, const double* coef, int64_t coef_batch_stride,
                    int64_t st1, int64_t st2, int64_t nbatch, void* stream = nullptr) const
    {
        deriv(0, 0, out, coords, npts, coef, coef_batch_stride, st1, st2, nbatch, stream);
    }
    // deriv_dim_1 / deriv_dim_2 / deriv_1_and_2 (:460-600) as derivative orders
    void deriv(int order1, int order2, double* out, const double* coords, int64_t npts, const double* coef,
               int64_t coef_batch_stride, int64_t st1, int64_t st2, int64_t nbatch, void* stream = nullptr) const
    {
        check(ddcx_evaluator2d_eval(e1_->handle(), e2_->handle(), order1, order2, out, coords, npts, coef,
                                    coef_batch_stride, st1, st2, nbatch, stream));
    }
    // bytes of device workspace the packed evaluation needs for nbatch
    // entries (it takes any amount from one entry's worth up; less than
    // nbatch entries' worth packs and evaluates a workspace-full at a time)
    int64_t eval_workspace(int64_t nbatch) const
    {
        int64_t bytes = 0;
        check(ddcx_evaluator2d_eval_workspace(e1_->handle(), e2_->handle(), nbatch, &bytes));
        return bytes;
    }
    // deriv (or operator() with orders 0, 0) through packed per-cell
    // coefficient records in `work` (16-byte aligned): for many points per
    // batch entry; the same values to rounding
    void deriv(int order1, int order2, double* out, const double* coords, int64_t npts, const double* coef,
               int64_t coef_batch_stride, int64_t st1, int64_t st2, int64_t nbatch, void* work, int64_t work_bytes,
               void* stream) const
    {
        check(ddcx_evaluator2d_eval_packed(e1_->handle(), e2_->handle(), order1, order2, out, coords, npts, coef,
                                           coef_batch_stride, st1, st2, nbatch, work, work_bytes, stream));
    }
    void deriv_dim_1(double* out, const double* coords, int64_t npts, const double* coef, int64_t cbs, int64_t st1,
                     int64_t st2, int64_t nbatch, void* stream = nullptr) const
    {
        deriv(1, 0, out, coords, npts, coef, cbs, st1, st2, nbatch, stream);
    }
    void deriv_dim_2(double* out, const double* coords, int64_t npts, const double* coef, int64_t cbs, int64_t st1,
                     int64_t st2, int64_t nbatch, void* stream = nullptr) const
    {
        deriv(0, 1, out, coords, npts, coef, cbs, st1, st2, nbatch, stream);
    }
    void deriv_1_and_2(double* out, const double* coords, int64_t npts, const double* coef, int64_t cbs,
                       int64_t st1, int64_t st2, int64_t nbatch, void* stream = nullptr) const
    {
        deriv(1, 1, out, coords, npts, coef, cbs, st1, st2, nbatch, stream);
    }
    // bytes of device workspace eval_vjp needs (0: no workspace)
    int64_t eval_vjp_workspace(int64_t npts, int64_t nbatch) const
    {
        int64_t bytes = 0;
        check(ddcx_evaluator2d_eval_vjp_workspace(e1_->handle(), e2_->handle(), npts, nbatch, &bytes));
        return bytes;
    }
    // adjoint of operator(): gcoef (strides of coef, overwritten) and gx
    // (interleaved like coords) from the weights w (nbatch*npts).  With a
    // workspace of eval_vjp_workspace() bytes the points are binned by cell
    // tile; with none, one atomic per point and coefficient.
    void eval_vjp(const double* coords, const double* w, double* gx, int64_t npts, const double* coef, double* gcoef,
                  int64_t coef_batch_stride, int64_t st1, int64_t st2, int64_t nbatch, void* work = nullptr,
                  int64_t work_bytes = 0, void* stream = nullptr) const
    {
        if (work && work_bytes > 0)
            check(ddcx_evaluator2d_eval_vjp_binned(e1_->handle(), e2_->handle(), coords, w, gx, npts, coef, gcoef,
                                                   coef_batch_stride, st1, st2, nbatch, work, work_bytes, stream));
        else
            check(ddcx_evaluator2d_eval_vjp(e1_->handle(), e2_->handle(), coords, w, gx, npts, coef, gcoef,
                                            coef_batch_stride, st1, st2, nbatch, stream));
    }

private:
    SplineEvaluator const* e1_;
    SplineEvaluator const* e2_;
};

// SplineEvaluatorND<Exec, Mem, TypeSeq<BSplines...>, TypeSeq<EvalDDim...>,
// TypeSeq<Rules...>> (spline_evaluator_nd.hpp:49): one 1D evaluator per
// dimension (1 to 4, Null or Periodic rules; not owned).  Points interleave
// their nd coordinates (point p of batch l at coords[nd*(l*npts + p)]);
// coefficients of batch l start at coef + l*coef_batch_stride, element
// (i_1..i_nd) at sum_k i_k*strides[k].
class SplineEvaluatorND {
public:
    explicit SplineEvaluatorND(std::vector<SplineEvaluator const*> evs)
    {
        for (auto const* e : evs) hs_.push_back(e->handle());
    }
    int dimension() const { return int(hs_.size()); }
    // operator()(spline_eval, coords_eval, spline_coef), spline_evaluator_nd.hpp:300-440
    void operator()(double* out, const double* coords, int64_t npts, const double* coef, int64_t coef_batch_stride,
                    std::vector<int64_t> const& strides, int64_t nbatch, void* stream = nullptr) const
    {
        deriv({}, out, coords, npts, coef, coef_batch_stride, strides, nbatch, stream);
    }
    // deriv(deriv_order, spline_eval, coords_eval, spline_coef), spline_evaluator_nd.hpp:445-590
    void deriv(std::vector<int> const& orders, double* out, const double* coords, int64_t npts, const double* coef,
               int64_t coef_batch_stride, std::vector<int64_t> const& strides, int64_t nbatch,
               void* stream = nullptr) const
    {
        if (int(strides.size()) != dimension() || (!orders.empty() && int(orders.size()) != dimension()))
            throw std::runtime_error("ddcx: one stride and one order per dimension");
        check(ddcx_evaluatornd_eval(hs_.data(), dimension(), orders.empty() ? nullptr : orders.data(), out, coords,
                                    npts, coef, coef_batch_stride, strides.data(), nbatch, stream));
    }
    // bytes of device workspace the binned evaluation needs for npts x
    // nbatch points (0: not available for this grid)
    int64_t eval_workspace(int64_t npts, int64_t nbatch) const
    {
        int64_t bytes = 0;
        check(ddcx_evaluatornd_eval_workspace(hs_.data(), dimension(), npts, nbatch, &bytes));
        return bytes;
    }
    // deriv (orders empty: operator()) with the points binned by tile of
    // cells through `work` (8-byte aligned): for many points per batch entry;
    // the same values bit for bit
    void deriv(std::vector<int> const& orders, double* out, const double* coords, int64_t npts, const double* coef,
               int64_t coef_batch_stride, std::vector<int64_t> const& strides, int64_t nbatch, void* work,
               int64_t work_bytes, void* stream) const
    {
        if (int(strides.size()) != dimension() || (!orders.empty() && int(orders.size()) != dimension()))
            throw std::runtime_error("ddcx: one stride and one order per dimension");
        check(ddcx_evaluatornd_eval_binned(hs_.data(), dimension(), orders.empty() ? nullptr : orders.data(), out,
                                           coords, npts, coef, coef_batch_stride, strides.data(), nbatch, work,
                                           work_bytes, stream));
    }
    // integrate(integrals, spline_coef), spline_evaluator_nd.hpp:597-632
    void integrate(double* out, int64_t out_stride, const double* coef, int64_t coef_batch_stride,
                   std::vector<int64_t> const& strides, int64_t nbatch, void* stream = nullptr) const
    {
        if (int(strides.size()) != dimension()) throw std::runtime_error("ddcx: one stride per dimension");
        check(ddcx_evaluatornd_integrate(hs_.data(), dimension(), out, out_stride, coef, coef_batch_stride,
                                         strides.data(), nbatch, stream));
    }
    // the same with a caller workspace of integrate_workspace_size(nbatch)
    // bytes, over which each batch entry's reduction is split
    int64_t integrate_workspace_size(int64_t nbatch) const
    {
        int64_t bytes = 0;
        check(ddcx_evaluatornd_integrate_workspace(hs_.data(), dimension(), nbatch, &bytes));
        return bytes;
    }
    void integrate(double* out, int64_t out_stride, const double* coef, int64_t coef_batch_stride,
                   std::vector<int64_t> const& strides, int64_t nbatch, void* work, int64_t work_bytes,
                   void* stream = nullptr) const
    {
        if (int(strides.size()) != dimension()) throw std::runtime_error("ddcx: one stride per dimension");
        check(ddcx_evaluatornd_integrate_ws(hs_.data(), dimension(), out, out_stride, coef, coef_batch_stride,
                                            strides.data(), nbatch, work, work_bytes, stream));
    }

private:
    std::vector<const ddcx_evaluator*> hs_;
};

// SplineEvaluator3D<...> (spline_evaluator_3d.hpp:56): the three-dimensional case.
class SplineEvaluator3D : public SplineEvaluatorND {
public:
    SplineEvaluator3D(SplineEvaluator const& e1, SplineEvaluator const& e2, SplineEvaluator const& e3)
        : SplineEvaluatorND({&e1, &e2, &e3})
    {
    }
};

// The derivative inputs of SplineBuilder3D::operator() with Hermite closures
// (spline_builder_3d.hpp:779-1090): layout_right device arrays named as the
// reference's arguments, nullptr where the closures take none.  With e_k =
// max(nbv_lower, nbv_upper) of dimension k's builder: derivs_*1 (nbatch, e1,
// n2, n3), derivs_*2 (nbatch, n1, e2, n3), derivs_*3 (nbatch, n1, n2, e3);
// mixed12[a][b] = mixed_derivs_{a}1_{b}2 (nbatch, e1, e2, n3), mixed23[a][b]
// = mixed_derivs_{a}2_{b}3 (nbatch, n1, e2, e3), mixed13[a][b] =
// mixed_derivs_{a}1_{b}3 (nbatch, e1, n2, e3), mixed123[a][b][c] =
// mixed_derivs_{a}1_{b}2_{c}3 (nbatch, e1, e2, e3); index 0 is min, 1 max.
// The gradients SplineBuilder3D::build_vjp returns with Hermite closures,
// one per derivative input of SplineBuilder3DDerivs and with the same extents
// (overwritten; nullptr where the closures take no such input).
struct SplineBuilder3DGrads {
    double* derivs_min1 = nullptr;
    double* derivs_max1 = nullptr;
    double* derivs_min2 = nullptr;
    double* derivs_max2 = nullptr;
    double* derivs_min3 = nullptr;
    double* derivs_max3 = nullptr;
    double* mixed12[2][2] = {};
    double* mixed23[2][2] = {};
    double* mixed13[2][2] = {};
    double* mixed123[2][2][2] = {};
};

struct SplineBuilder3DDerivs {
    const double* derivs_min1 = nullptr;
    const double* derivs_max1 = nullptr;
    const double* derivs_min2 = nullptr;
    const double* derivs_max2 = nullptr;
    const double* derivs_min3 = nullptr;
    const double* derivs_max3 = nullptr;
    const double* mixed12[2][2] = {};
    const double* mixed23[2][2] = {};
    const double* mixed13[2][2] = {};
    const double* mixed123[2][2][2] = {};
};

// SplineBuilder3D<...> (spline_builder_3d.hpp:45): one batched 1D build per
// dimension, dimension 1 first, on layout_right arrays vals (nbatch, n1, n2,
// n3) and spline (nbatch, s1, s2, s3), through two caller-provided device
// workspaces of work1_size() and work2_size() doubles (nothing is allocated).
// With Hermite closures the call takes the derivative inputs and one
// workspace of hermite_workspace_size() doubles, and composes the builds as
// the reference does: the dimension-2 and dimension-3 derivatives and the
// 2-3 cross derivatives are built along dimension 1 (with the mixed
// derivatives involving dimension 1 as their Hermite data), the values along
// dimension 1, then SplineBuilder2D over dimensions 2 and 3.
class SplineBuilder3D {
public:
    SplineBuilder3D(SplineBuilder const& b1, SplineBuilder const& b2, SplineBuilder const& b3, int64_t size1,
                    int64_t size2, int64_t size3)
        : b_{&b1, &b2, &b3}, s_{size1, size2, size3}
    {
    }
    int64_t work1_size(int64_t nbatch) const { return nbatch * s_[0] * n(1) * n(2); }
    int64_t work2_size(int64_t nbatch) const { return nbatch * s_[0] * s_[1] * n(2); }
    // closures without derivative inputs
    void operator()(double* spline, const double* vals, int64_t nbatch, double* work1, double* work2,
                    void* stream = nullptr) const
    {
        if (hermite()) throw std::runtime_error("ddcx: Hermite closures take the derivative inputs");
        pass(0, work1, {nbatch, s_[0], n(1), n(2)}, vals, {nbatch, n(0), n(1), n(2)}, stream);
        pass(1, work2, {nbatch, s_[0], s_[1], n(2)}, work1, {nbatch, s_[0], n(1), n(2)}, stream);
        pass(2, spline, {nbatch, s_[0], s_[1], s_[2]}, work2, {nbatch, s_[0], s_[1], n(2)}, stream);
    }
    int64_t hermite_workspace_size(int64_t nbatch) const
    {
        int64_t const e2 = e(1), e3 = e(2), nb1 = nbatch * s_[0];
        int64_t const w = nb1 * n(1) * n(2) + 2 * nb1 * e2 * n(2) + 2 * nb1 * n(1) * e3 + 4 * nb1 * e2 * e3;
        return w + SplineBuilder2D(*b_[1], *b_[2], s_[1]).workspace_size(nb1);
    }
    // any closures, with the derivative inputs they take
    void operator()(double* spline, const double* vals, int64_t nbatch, double* work, SplineBuilder3DDerivs const& d,
                    void* stream = nullptr) const
    {
        int64_t const n1 = n(0), n2 = n(1), n3 = n(2), s1 = s_[0], e1 = e(0), e2 = e(1), e3 = e(2);
        bool const h2[2] = {b_[1]->info().nbv_lower > 0, b_[1]->info().nbv_upper > 0};
        bool const h3[2] = {b_[2]->info().nbv_lower > 0, b_[2]->info().nbv_upper > 0};
        double* p = work;
        auto take = [&](int64_t count) {
            double* q = p;
            p += count;
            return q;
        };
        double* spline1 = take(nbatch * s1 * n2 * n3);
        double* d2[2] = {take(nbatch * s1 * e2 * n3), take(nbatch * s1 * e2 * n3)};
        double* d3[2] = {take(nbatch * s1 * n2 * e3), take(nbatch * s1 * n2 * e3)};
        double* m23[2][2];
        for (auto& row : m23)
            for (auto& m : row) m = take(nbatch * s1 * e2 * e3);
        // one dimension-1 build: in (nbatch, n1, a, b) -> out (nbatch, s1, a, b)
        auto along1 = [&](double* out, const double* in, int64_t a, int64_t b, const double* lo, const double* hi) {
            if (!in) throw std::runtime_error("ddcx: a derivative input the Hermite closures take is missing");
            int64_t nb = 0;
            ddcx_layout const l_out = batched_layout({nbatch, s1, a, b}, 1, &nb);
            ddcx_layout const l_in = batched_layout({nbatch, n1, a, b}, 1);
            ddcx_layout const l_d = batched_layout({nbatch, std::max<int64_t>(e1, 1), a, b}, 1);
            check(ddcx_builder_build(b_[0]->handle(), out, l_out, in, l_in, lo, hi, l_d, nb, stream));
        };
        const double* dk2[2] = {d.derivs_min2, d.derivs_max2};
        const double* dk3[2] = {d.derivs_min3, d.derivs_max3};
        for (int sd = 0; sd < 2; ++sd)
            if (h2[sd]) along1(d2[sd], dk2[sd], e2, n3, d.mixed12[0][sd], d.mixed12[1][sd]);
        for (int sd = 0; sd < 2; ++sd)
            if (h3[sd]) along1(d3[sd], dk3[sd], n2, e3, d.mixed13[0][sd], d.mixed13[1][sd]);
        for (int a2 = 0; a2 < 2; ++a2)
            for (int a3 = 0; a3 < 2; ++a3)
                if (h2[a2] && h3[a3])
                    along1(m23[a2][a3], d.mixed23[a2][a3], e2, e3, d.mixed123[0][a2][a3], d.mixed123[1][a2][a3]);
        along1(spline1, vals, n2, n3, d.derivs_min1, d.derivs_max1);
        auto opt = [](bool on, double* q) { return on ? static_cast<const double*>(q) : nullptr; };
        SplineBuilder2D(*b_[1], *b_[2], s_[1])(spline, spline1, nbatch * s1, p, opt(h2[0], d2[0]), opt(h2[1], d2[1]),
                                               opt(h3[0], d3[0]), opt(h3[1], d3[1]),
                                               opt(h2[0] && h3[0], m23[0][0]), opt(h2[1] && h3[0], m23[1][0]),
                                               opt(h2[0] && h3[1], m23[0][1]), opt(h2[1] && h3[1], m23[1][1]),
                                               stream);
    }

    // adjoint of the form without derivative inputs: the three transposed
    // passes in reverse order, through the same two workspaces
    void build_vjp(const double* gspline, double* gvals, int64_t nbatch, double* work1, double* work2,
                   void* stream = nullptr) const
    {
        if (hermite()) throw std::runtime_error("ddcx: Hermite closures return the derivative gradients");
        pass_vjp(2, gspline, {nbatch, s_[0], s_[1], s_[2]}, work2, {nbatch, s_[0], s_[1], n(2)}, stream);
        pass_vjp(1, work2, {nbatch, s_[0], s_[1], n(2)}, work1, {nbatch, s_[0], n(1), n(2)}, stream);
        pass_vjp(0, work1, {nbatch, s_[0], n(1), n(2)}, gvals, {nbatch, n(0), n(1), n(2)}, stream);
    }
    // adjoint of the Hermite form: gvals and the derivative gradients (every
    // one the closures take must be given), through a workspace of
    // hermite_workspace_size() doubles; operator()'s builds in reverse order
    void build_vjp(const double* gspline, double* gvals, int64_t nbatch, double* work, SplineBuilder3DGrads const& g,
                   void* stream = nullptr) const
    {
        int64_t const n1 = n(0), n2 = n(1), n3 = n(2), s1 = s_[0], e1 = e(0), e2 = e(1), e3 = e(2);
        bool const h2[2] = {b_[1]->info().nbv_lower > 0, b_[1]->info().nbv_upper > 0};
        bool const h3[2] = {b_[2]->info().nbv_lower > 0, b_[2]->info().nbv_upper > 0};
        double* p = work;
        auto take = [&](int64_t count) {
            double* q = p;
            p += count;
            return q;
        };
        double* gs1 = take(nbatch * s1 * n2 * n3);
        double* gd2[2] = {take(nbatch * s1 * e2 * n3), take(nbatch * s1 * e2 * n3)};
        double* gd3[2] = {take(nbatch * s1 * n2 * e3), take(nbatch * s1 * n2 * e3)};
        double* gm23[2][2];
        for (auto& row : gm23)
            for (auto& m : row) m = take(nbatch * s1 * e2 * e3);
        auto opt = [](bool on, double* q) { return on ? q : nullptr; };
        SplineBuilder2D(*b_[1], *b_[2], s_[1]).build_vjp(gspline, gs1, nbatch * s1, p, opt(h2[0], gd2[0]),
                                                         opt(h2[1], gd2[1]), opt(h3[0], gd3[0]), opt(h3[1], gd3[1]),
                                                         opt(h2[0] && h3[0], gm23[0][0]),
                                                         opt(h2[1] && h3[0], gm23[1][0]),
                                                         opt(h2[0] && h3[1], gm23[0][1]),
                                                         opt(h2[1] && h3[1], gm23[1][1]), stream);
        // one transposed dimension-1 build: gin (nbatch, s1, a, b) -> gout (nbatch, n1, a, b)
        auto along1 = [&](const double* gin, double* gout, int64_t a, int64_t b, double* glo, double* ghi) {
            if (!gout) throw std::runtime_error("ddcx: a gradient of a derivative input the closures take is missing");
            int64_t nb = 0;
            ddcx_layout const l_in = batched_layout({nbatch, s1, a, b}, 1, &nb);
            ddcx_layout const l_out = batched_layout({nbatch, n1, a, b}, 1);
            ddcx_layout const l_d = batched_layout({nbatch, std::max<int64_t>(e1, 1), a, b}, 1);
            check(ddcx_builder_build_vjp(b_[0]->handle(), gin, l_in, gout, l_out, glo, ghi, l_d, nb, stream));
        };
        along1(gs1, gvals, n2, n3, g.derivs_min1, g.derivs_max1);
        double* gk2[2] = {g.derivs_min2, g.derivs_max2};
        double* gk3[2] = {g.derivs_min3, g.derivs_max3};
        for (int sd = 0; sd < 2; ++sd)
            if (h2[sd]) along1(gd2[sd], gk2[sd], e2, n3, g.mixed12[0][sd], g.mixed12[1][sd]);
        for (int sd = 0; sd < 2; ++sd)
            if (h3[sd]) along1(gd3[sd], gk3[sd], n2, e3, g.mixed13[0][sd], g.mixed13[1][sd]);
        for (int a2 = 0; a2 < 2; ++a2)
            for (int a3 = 0; a3 < 2; ++a3)
                if (h2[a2] && h3[a3])
                    along1(gm23[a2][a3], g.mixed23[a2][a3], e2, e3, g.mixed123[0][a2][a3], g.mixed123[1][a2][a3]);
    }

private:
    int64_t n(int k) const { return b_[k]->npoints(); }
    int64_t e(int k) const { return std::max(b_[k]->info().nbv_lower, b_[k]->info().nbv_upper); }
    void pass_vjp(int k, const double* gin, std::vector<int64_t> const& ei, double* gout,
                  std::vector<int64_t> const& eo, void* stream) const
    {
        int64_t nb = 0;
        ddcx_layout const li = batched_layout(ei, k + 1, &nb);
        ddcx_layout const lo = batched_layout(eo, k + 1);
        b_[k]->build_vjp(Span<double const>{gin, li, nb}, Span<double>{gout, lo, nb}, std::nullopt, std::nullopt,
                         stream);
    }
    bool hermite() const
    {
        for (auto const* b : b_)
            if (b->info().nbv_lower || b->info().nbv_upper) return true;
        return false;
    }
    void pass(int k, double* out, std::vector<int64_t> const& eo, const double* in, std::vector<int64_t> const& ei,
              void* stream) const
    {
        int64_t nb = 0;
        ddcx_layout const lo = batched_layout(eo, k + 1, &nb);
        ddcx_layout const li = batched_layout(ei, k + 1);
        (*b_[k])(Span<double>{out, lo, nb}, Span<double const>{in, li, nb}, std::nullopt, std::nullopt, stream);
    }
    SplineBuilder const* b_[3];
    int64_t s_[3];
};

} // namespace ddcx
