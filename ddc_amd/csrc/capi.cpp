// extern "C" boundary of libddcx (declared in include/ddcx.h).  Converts C
// handles and layouts to the internal plans, catches every exception and
// returns a status with a thread-local message.
#include <algorithm>
#include <cmath>
#include <memory>
#include <cstring>
#include <string>

#include "internal.hpp"

struct ddcx_bsplines {
    std::shared_ptr<ddcx::BSplinesPlan> p;
};
struct ddcx_builder {
    std::unique_ptr<ddcx::BuilderPlan> p;
};
struct ddcx_evaluator {
    std::shared_ptr<ddcx::BSplinesPlan> bs;
    ddcx_extrap lower, upper;
    // Constant rules: ConstantExtrapolationRule's eval_pos (resolved) and,
    // for 2D use, the clamp range of the other coordinate (NaN: that
    // dimension's [rmin, rmax])
    double pos[2];
    double ni_min[2], ni_max[2];
};
struct ddcx_problem {
    ddcx::ProblemPlan p;
};

namespace {

thread_local std::string g_last_error;

template <class F>
ddcx_status guarded(F&& f)
{
    try {
        f();
        g_last_error.clear();
        return DDCX_OK;
    } catch (ddcx::Error const& e) {
        g_last_error = e.what();
        return e.status;
    } catch (std::exception const& e) {
        g_last_error = e.what();
        return DDCX_ERR_INVALID;
    } catch (...) {
        g_last_error = "unknown error";
        return DDCX_ERR_INVALID;
    }
}

void require(bool cond, const char* msg)
{
    if (!cond) throw ddcx::Error(DDCX_ERR_INVALID, msg);
}

int rule_of(ddcx_extrap e)
{
    return e == DDCX_EXTRAP_CONSTANT ? DDCX_EXTRAP_CONSTANT : DDCX_EXTRAP_NULL;
}

ddcx::Rules1D rules1d(const ddcx_evaluator* h)
{
    return ddcx::Rules1D{rule_of(h->lower), rule_of(h->upper), h->pos[0], h->pos[1]};
}

// The four rules of SplineEvaluator2D(e1's lower, upper, e2's lower, upper);
// a Constant rule's default clamp range is the other dimension's domain.
ddcx::Rules2D rules2d(const ddcx_evaluator* e1, const ddcx_evaluator* e2)
{
    ddcx::Rules2D r{};
    const ddcx_evaluator* ev[2] = {e1, e2};
    for (int d = 0; d < 2; ++d) {
        ddcx::BSplinesView const& other = ev[1 - d]->bs->view;
        for (int s = 0; s < 2; ++s) {
            int const k = 2 * d + s;
            r.kind[k] = rule_of(s == 0 ? ev[d]->lower : ev[d]->upper);
            r.pos[k] = ev[d]->pos[s];
            r.ni_min[k] = std::isnan(ev[d]->ni_min[s]) ? other.rmin : ev[d]->ni_min[s];
            r.ni_max[k] = std::isnan(ev[d]->ni_max[s]) ? other.rmax : ev[d]->ni_max[s];
            require(r.ni_min[k] <= r.ni_max[k],
                    "ConstantExtrapolationRule: eval_pos_not_interest_min exceeds eval_pos_not_interest_max");
        }
    }
    return r;
}

// ddcx_evaluator_create_with_rules: one rule's positions, checked
void set_rule(ddcx_evaluator* h, int side, ddcx_extrap_rule const& r)
{
    ddcx::BSplinesView const& v = h->bs->view;
    (side == 0 ? h->lower : h->upper) = r.kind;
    double const bound = side == 0 ? v.rmin : v.rmax;
    h->pos[side] = bound;
    h->ni_min[side] = h->ni_max[side] = std::nan("");
    if (r.kind != DDCX_EXTRAP_CONSTANT) return;
    if (!std::isnan(r.eval_pos)) {
        require(r.eval_pos >= v.rmin && r.eval_pos <= v.rmax,
                "ConstantExtrapolationRule: eval_pos must lie inside the domain [rmin, rmax]");
        h->pos[side] = r.eval_pos;
    }
    require(std::isnan(r.eval_pos_not_interest_min) == std::isnan(r.eval_pos_not_interest_max),
            "ConstantExtrapolationRule: give both eval_pos_not_interest_min and _max, or neither");
    h->ni_min[side] = r.eval_pos_not_interest_min;
    h->ni_max[side] = r.eval_pos_not_interest_max;
}

} // namespace

extern "C" {

const char* ddcx_last_error(void) { return g_last_error.c_str(); }
const char* ddcx_version(void) { return "ddcx 0.1.0 (gfx950)"; }

ddcx_status ddcx_bsplines_create(const ddcx_bsplines_desc* desc, ddcx_bsplines** out)
{
    return guarded([&] {
        require(desc && out, "null argument");
        *out = nullptr;
        auto h = new ddcx_bsplines{ddcx::make_bsplines(*desc)};
        *out = h;
    });
}

ddcx_status ddcx_bsplines_destroy(ddcx_bsplines* b)
{
    return guarded([&] { delete b; });
}

ddcx_status ddcx_bsplines_knots(const ddcx_bsplines* b, double* knots)
{
    return guarded([&] {
        require(b && knots, "null argument");
        std::memcpy(knots, b->p->knots_host.data(), b->p->knots_host.size() * sizeof(double));
    });
}

ddcx_status ddcx_bsplines_integrals(const ddcx_bsplines* b, double* out)
{
    return guarded([&] {
        require(b && out, "null argument");
        std::memcpy(out, b->p->integrals.data(), b->p->integrals.size() * sizeof(double));
    });
}

ddcx_status ddcx_builder_create(const ddcx_bsplines* b, const ddcx_builder_desc* desc, ddcx_builder** out)
{
    return guarded([&] {
        require(b && desc && out, "null argument");
        *out = nullptr;
        auto h = new ddcx_builder{ddcx::make_builder(b->p, *desc)};
        *out = h;
    });
}

ddcx_status ddcx_builder_destroy(ddcx_builder* h)
{
    return guarded([&] { delete h; });
}

ddcx_status ddcx_builder_get_info(const ddcx_builder* h, ddcx_builder_info* info)
{
    return guarded([&] {
        require(h && info, "null argument");
        ddcx::BuilderPlan const& p = *h->p;
        info->npoints = int(p.points.size());
        info->nbasis = p.bs->nbasis();
        info->size = p.bs->size();
        info->offset = p.offset;
        info->nbv_lower = p.nbv_l;
        info->nbv_upper = p.nbv_r;
        info->matrix_kind = int(p.prob.q.type);
        info->blocked = p.prob.blocked ? (p.prob.top > 0 ? 2 : 1) : 0;
        info->main_size = p.prob.q.n;
        info->border = p.prob.nb();
        info->dx = p.dx;
        info->kernel = p.refine.ok ? 2 : p.tiled.ok ? 1 : 0;
        info->refine_steps = p.refine.ok ? p.refine.steps : 0;
        info->refine_contraction = p.refine.ok ? p.refine.rho : 0.0;
        info->pivoted = 0;
        for (size_t j = 0; j < p.prob.q.ipiv.size(); ++j)
            if (p.prob.q.ipiv[j] != int(j) && p.prob.q.type == ddcx::QType::Band) info->pivoted = 1;
    });
}

ddcx_status ddcx_builder_points(const ddcx_builder* h, double* points)
{
    return guarded([&] {
        require(h && points, "null argument");
        std::memcpy(points, h->p->points.data(), h->p->points.size() * sizeof(double));
    });
}

ddcx_status ddcx_builder_build(const ddcx_builder* h, double* coef, ddcx_layout coef_layout, const double* vals,
                               ddcx_layout vals_layout, const double* derivs_xmin, const double* derivs_xmax,
                               ddcx_layout derivs_layout, int64_t nbatch, void* stream)
{
    return guarded([&] {
        require(h, "null builder");
        require(nbatch >= 0, "negative batch size");
        if (nbatch == 0) return; // spline_builder.hpp:929-931
        require(coef && vals, "null data pointer");
        ddcx::BuilderPlan const& p = *h->p;
        if (p.bcl == DDCX_HERMITE && p.nbv_l > 0) require(derivs_xmin != nullptr, "HERMITE lower closure needs derivs_xmin");
        if (p.bcu == DDCX_HERMITE && p.nbv_r > 0) require(derivs_xmax != nullptr, "HERMITE upper closure needs derivs_xmax");
        ddcx::launch_build(p, coef, ddcx::to_layout(coef_layout), vals, ddcx::to_layout(vals_layout), derivs_xmin,
                           derivs_xmax, ddcx::to_layout(derivs_layout), nbatch, static_cast<hipStream_t>(stream));
    });
}

ddcx_status ddcx_build_eval(const ddcx_builder* h, const ddcx_evaluator* e, double* coef, ddcx_layout coef_layout,
                            const double* vals, ddcx_layout vals_layout, const double* derivs_xmin,
                            const double* derivs_xmax, ddcx_layout derivs_layout, double* out, const double* coords,
                            ddcx_layout points_layout, int64_t npts, int64_t nbatch, void* stream)
{
    return guarded([&] {
        require(h && e, "null handle");
        require(npts >= 0 && nbatch >= 0, "negative size");
        if (nbatch == 0) return;
        require(vals && (npts == 0 || (out && coords)), "null data pointer");
        ddcx::BuilderPlan const& p = *h->p;
        if (p.bcl == DDCX_HERMITE && p.nbv_l > 0) require(derivs_xmin != nullptr, "HERMITE lower closure needs derivs_xmin");
        if (p.bcu == DDCX_HERMITE && p.nbv_r > 0) require(derivs_xmax != nullptr, "HERMITE upper closure needs derivs_xmax");
        require(p.bs.get() == e->bs.get(), "builder and evaluator must share one B-spline basis");
        auto const cl = ddcx::to_layout(coef_layout), vl = ddcx::to_layout(vals_layout),
                   dl = ddcx::to_layout(derivs_layout), pl = ddcx::to_layout(points_layout);
        auto const st = static_cast<hipStream_t>(stream);
        p.ensure_device();
        e->bs->ensure_device();
        if (p.precision == DDCX_FP64 && npts > 0 &&
            ddcx::launch_build_eval_fused(p, *e->bs, rules1d(e), coef, cl, vals, vl,
                                          derivs_xmin, derivs_xmax, dl, out, coords, pl, npts, nbatch, st))
            return;
        require(coef != nullptr,
                "the fused kernel does not cover this case (precision, layout or basis); pass a coefficient array");
        ddcx::launch_build(p, coef, cl, vals, vl, derivs_xmin, derivs_xmax, dl, nbatch, st);
        if (npts > 0)
            ddcx::launch_eval(*e->bs, rules1d(e), 0, out, coords, pl, npts, coef, cl, nbatch,
                              st);
    });
}

ddcx_status ddcx_advect(const ddcx_builder* h, const ddcx_evaluator* e, double* coef, ddcx_layout coef_layout,
                        const double* vals, ddcx_layout vals_layout, const double* derivs_xmin,
                        const double* derivs_xmax, ddcx_layout derivs_layout, const double* shift, double* out,
                        ddcx_layout out_layout, int64_t nbatch, void* stream)
{
    return guarded([&] {
        require(h && e, "null handle");
        require(nbatch >= 0, "negative size");
        if (nbatch == 0) return;
        require(vals && shift && out, "null data pointer");
        ddcx::BuilderPlan const& p = *h->p;
        if (p.bcl == DDCX_HERMITE && p.nbv_l > 0) require(derivs_xmin != nullptr, "HERMITE lower closure needs derivs_xmin");
        if (p.bcu == DDCX_HERMITE && p.nbv_r > 0) require(derivs_xmax != nullptr, "HERMITE upper closure needs derivs_xmax");
        require(p.bs.get() == e->bs.get(), "builder and evaluator must share one B-spline basis");
        auto const cl = ddcx::to_layout(coef_layout), vl = ddcx::to_layout(vals_layout),
                   dl = ddcx::to_layout(derivs_layout), ol = ddcx::to_layout(out_layout);
        p.ensure_device();
        e->bs->ensure_device();
        bool const done =
            p.precision == DDCX_FP64 &&
            ddcx::launch_build_eval_fused(p, *e->bs, rules1d(e), coef, cl, vals, vl,
                                          derivs_xmin, derivs_xmax, dl, out, nullptr, ol, int64_t(p.points.size()),
                                          nbatch, static_cast<hipStream_t>(stream), shift);
        if (!done)
            throw ddcx::Error(DDCX_ERR_UNSUPPORTED,
                              "advection step: the fused kernel does not cover this case (precision, layout or basis)");
    });
}

ddcx_status ddcx_builder_build_vjp(const ddcx_builder* h, const double* gcoef, ddcx_layout coef_layout, double* gvals,
                                   ddcx_layout vals_layout, double* gderivs_xmin, double* gderivs_xmax,
                                   ddcx_layout derivs_layout, int64_t nbatch, void* stream)
{
    return guarded([&] {
        require(h, "null builder");
        require(nbatch >= 0, "negative batch size");
        if (nbatch == 0) return;
        require(gcoef && gvals, "null data pointer");
        ddcx::BuilderPlan const& p = *h->p;
        if (p.bcl == DDCX_HERMITE && p.nbv_l > 0) require(gderivs_xmin != nullptr, "HERMITE lower closure needs gderivs_xmin");
        if (p.bcu == DDCX_HERMITE && p.nbv_r > 0) require(gderivs_xmax != nullptr, "HERMITE upper closure needs gderivs_xmax");
        ddcx::launch_build_vjp(p, gcoef, ddcx::to_layout(coef_layout), gvals, ddcx::to_layout(vals_layout),
                               gderivs_xmin, gderivs_xmax, ddcx::to_layout(derivs_layout), nbatch,
                               static_cast<hipStream_t>(stream));
    });
}

ddcx_status ddcx_builder_quadrature(const ddcx_builder* h, double* q_xmin, double* q_vals, double* q_xmax)
{
    return guarded([&] {
        require(h && q_vals, "null argument");
        ddcx::BuilderPlan const& p = *h->p;
        ddcx::HostProblem const& pr = p.prob;
        // spline_builder.hpp:1111-1234: solve A^T q = integrals (first size() rows)
        std::vector<double> m(pr.rhs_rows(), 0.0);
        for (int i = 0; i < pr.n; ++i) m[i] = p.bs->integrals[i];
        pr.solve(m.data(), 1, true);
        int const nv = p.bs->nbasis() - p.nbv_l - p.nbv_r;
        auto dxp = [&](int k) {
            double r = 1.0;
            for (int j = 0; j < k; ++j) r *= p.dx;
            return r;
        };
        for (int i = 0; i < p.nbv_l; ++i) q_xmin[i] = m[i] * dxp(i + p.odd);
        for (int i = 0; i < nv; ++i) q_vals[i] = m[p.nbv_l + i];
        for (int i = 0; i < p.nbv_r; ++i) q_xmax[i] = m[p.nbv_l + nv + i] * dxp(i + p.odd);
    });
}

ddcx_status ddcx_evaluator_create(const ddcx_bsplines* b, ddcx_extrap lower, ddcx_extrap upper, ddcx_evaluator** out)
{
    return guarded([&] {
        require(b && out, "null argument");
        *out = nullptr;
        ddcx_extrap_rule const nan_rule{DDCX_EXTRAP_NULL, std::nan(""), std::nan(""), std::nan("")};
        ddcx_extrap_rule lo = nan_rule, up = nan_rule;
        lo.kind = lower;
        up.kind = upper;
        auto h = std::make_unique<ddcx_evaluator>();
        h->bs = b->p;
        set_rule(h.get(), 0, lo);
        set_rule(h.get(), 1, up);
        *out = h.release();
    });
}

ddcx_status ddcx_evaluator_create_with_rules(const ddcx_bsplines* b, const ddcx_extrap_rule* lower,
                                             const ddcx_extrap_rule* upper, ddcx_evaluator** out)
{
    return guarded([&] {
        require(b && lower && upper && out, "null argument");
        *out = nullptr;
        auto h = std::make_unique<ddcx_evaluator>();
        h->bs = b->p;
        set_rule(h.get(), 0, *lower);
        set_rule(h.get(), 1, *upper);
        *out = h.release();
    });
}

ddcx_status ddcx_evaluator_destroy(ddcx_evaluator* h)
{
    return guarded([&] { delete h; });
}

ddcx_status ddcx_evaluator_eval(const ddcx_evaluator* h, double* out, const double* coords, ddcx_layout points_layout,
                                int64_t npts, const double* coef, ddcx_layout coef_layout, int64_t nbatch, void* stream)
{
    return guarded([&] {
        require(h, "null evaluator");
        require(npts >= 0 && nbatch >= 0, "negative size");
        if (npts * nbatch == 0) return;
        require(out && coords && coef, "null data pointer");
        ddcx::launch_eval(*h->bs, rules1d(h), 0, out, coords, ddcx::to_layout(points_layout),
                          npts, coef, ddcx::to_layout(coef_layout), nbatch, static_cast<hipStream_t>(stream));
    });
}

ddcx_status ddcx_evaluator_eval_at(const ddcx_evaluator* h, int order, double* out, ddcx_layout points_layout,
                                   const double* points, int64_t npts, const double* coef, ddcx_layout coef_layout,
                                   int64_t nbatch, void* stream)
{
    return guarded([&] {
        require(h, "null evaluator");
        require(order >= 0 && order <= h->bs->view.degree, "derivative order must be in [0, degree]");
        require(npts >= 0 && nbatch >= 0, "negative size");
        if (npts * nbatch == 0) return;
        require(out && points && coef, "null data pointer");
        ddcx::Layout const shared{1, 1, 0, 0}; // point p of every line at points[p]
        ddcx::launch_eval_xl(*h->bs, rules1d(h), order, out, points, shared, ddcx::to_layout(points_layout), npts,
                             coef, ddcx::to_layout(coef_layout), nbatch, static_cast<hipStream_t>(stream));
    });
}

ddcx_status ddcx_evaluator_deriv(const ddcx_evaluator* h, int order, double* out, const double* coords,
                                 ddcx_layout points_layout, int64_t npts, const double* coef, ddcx_layout coef_layout,
                                 int64_t nbatch, void* stream)
{
    return guarded([&] {
        require(h, "null evaluator");
        require(order > 0 && order <= h->bs->view.degree, "derivative order must be in [1, degree]");
        require(npts >= 0 && nbatch >= 0, "negative size");
        if (npts * nbatch == 0) return;
        require(out && coords && coef, "null data pointer");
        ddcx::launch_eval(*h->bs, rules1d(h), order, out, coords,
                          ddcx::to_layout(points_layout), npts, coef, ddcx::to_layout(coef_layout), nbatch,
                          static_cast<hipStream_t>(stream));
    });
}

ddcx_status ddcx_evaluator_device_view(const ddcx_evaluator* h, ddcx_device_evaluator* out)
{
    return guarded([&] {
        require(h && out, "null argument");
        h->bs->ensure_device();
        ddcx::BSplinesView const v = h->bs->device_view();
        ddcx::Rules1D const r = rules1d(h);
        *out = ddcx_device_evaluator{v.degree, v.uniform, v.periodic, v.ncells, v.origin, v.step, v.inv_step,
                                     v.rmin, v.rmax, v.length, v.knots, v.cells, v.nbucket, v.inv_bucket,
                                     r.lower, r.upper, r.pos_lower, r.pos_upper};
    });
}

ddcx_status ddcx_evaluator_integrate(const ddcx_evaluator* h, double* out, int64_t out_stride, const double* coef,
                                     ddcx_layout coef_layout, int64_t nbatch, void* stream)
{
    return guarded([&] {
        require(h, "null evaluator");
        require(nbatch >= 0, "negative size");
        if (nbatch == 0) return;
        require(out && coef, "null data pointer");
        ddcx::launch_integrate(*h->bs, out, out_stride, coef, ddcx::to_layout(coef_layout), nbatch,
                               static_cast<hipStream_t>(stream));
    });
}

ddcx_status ddcx_evaluator_eval_vjp(const ddcx_evaluator* h, const double* coords, const double* w, double* gx,
                                    ddcx_layout points_layout, int64_t npts, const double* coef, double* gcoef,
                                    ddcx_layout coef_layout, int64_t nbatch, void* stream)
{
    return guarded([&] {
        require(h, "null evaluator");
        require(npts >= 0 && nbatch >= 0, "negative size");
        if (nbatch == 0) return;
        require(gcoef && coef, "null data pointer");
        if (npts > 0) require(coords && w && gx, "null data pointer");
        ddcx::launch_eval_vjp(*h->bs, rules1d(h), coords, w, gx,
                              ddcx::to_layout(points_layout), npts, coef, gcoef, ddcx::to_layout(coef_layout), nbatch,
                              static_cast<hipStream_t>(stream));
    });
}

ddcx_status ddcx_evaluator2d_eval(const ddcx_evaluator* e1, const ddcx_evaluator* e2, int order1, int order2,
                                  double* out, const double* coords, int64_t npts, const double* coef,
                                  int64_t coef_batch_stride, int64_t st1, int64_t st2, int64_t nbatch, void* stream)
{
    return guarded([&] {
        require(e1 && e2, "null evaluator");
        require(order1 >= 0 && order2 >= 0, "negative derivative order");
        require(npts >= 0 && nbatch >= 0, "negative size");
        if (npts * nbatch == 0) return;
        require(out && coords && coef, "null data pointer");
        ddcx::Rules2D const rules = rules2d(e1, e2);
        ddcx::launch_eval2d(*e1->bs, *e2->bs, rules, order1, order2, out, coords, npts, coef, coef_batch_stride, st1,
                            st2, nbatch, static_cast<hipStream_t>(stream));
    });
}

ddcx_status ddcx_evaluator2d_eval_workspace(const ddcx_evaluator* e1, const ddcx_evaluator* e2, int64_t nbatch,
                                            int64_t* bytes)
{
    return guarded([&] {
        require(e1 && e2 && bytes, "null argument");
        require(nbatch >= 0, "negative size");
        *bytes = ddcx::eval2d_packed_record_bytes(*e1->bs, *e2->bs) * nbatch;
    });
}

ddcx_status ddcx_evaluator2d_eval_packed(const ddcx_evaluator* e1, const ddcx_evaluator* e2, int order1, int order2,
                                         double* out, const double* coords, int64_t npts, const double* coef,
                                         int64_t coef_batch_stride, int64_t st1, int64_t st2, int64_t nbatch,
                                         void* work, int64_t work_bytes, void* stream)
{
    return guarded([&] {
        require(e1 && e2, "null evaluator");
        require(order1 >= 0 && order2 >= 0, "negative derivative order");
        require(npts >= 0 && nbatch >= 0, "negative size");
        if (npts * nbatch == 0) return;
        require(out && coords && coef && work, "null data pointer");
        ddcx::Rules2D const rules = rules2d(e1, e2);
        ddcx::launch_eval2d_packed(*e1->bs, *e2->bs, rules, order1, order2, out, coords, npts, coef,
                                   coef_batch_stride, st1, st2, nbatch, work, work_bytes,
                                   static_cast<hipStream_t>(stream));
    });
}

ddcx_status ddcx_evaluator2d_eval_vjp(const ddcx_evaluator* e1, const ddcx_evaluator* e2, const double* coords,
                                      const double* w, double* gx, int64_t npts, const double* coef, double* gcoef,
                                      int64_t coef_batch_stride, int64_t st1, int64_t st2, int64_t nbatch,
                                      void* stream)
{
    return guarded([&] {
        require(e1 && e2, "null evaluator");
        require(npts >= 0 && nbatch >= 0, "negative size");
        if (nbatch == 0) return;
        require(coef && gcoef, "null data pointer");
        require(npts == 0 || (coords && w && gx), "null data pointer");
        ddcx::Rules2D const rules = rules2d(e1, e2);
        ddcx::launch_eval2d_vjp(*e1->bs, *e2->bs, rules, coords, w, gx, npts, coef, gcoef, coef_batch_stride, st1,
                                st2, nbatch, static_cast<hipStream_t>(stream));
    });
}

ddcx_status ddcx_evaluator2d_eval_vjp_workspace(const ddcx_evaluator* e1, const ddcx_evaluator* e2, int64_t npts,
                                                int64_t nbatch, int64_t* bytes)
{
    return guarded([&] {
        require(e1 && e2 && bytes, "null argument");
        require(npts >= 0 && nbatch >= 0, "negative size");
        *bytes = ddcx::eval2d_vjp_workspace(*e1->bs, *e2->bs, npts, nbatch);
    });
}

ddcx_status ddcx_evaluator2d_eval_vjp_binned(const ddcx_evaluator* e1, const ddcx_evaluator* e2,
                                             const double* coords, const double* w, double* gx, int64_t npts,
                                             const double* coef, double* gcoef, int64_t coef_batch_stride,
                                             int64_t st1, int64_t st2, int64_t nbatch, void* work,
                                             int64_t work_bytes, void* stream)
{
    return guarded([&] {
        require(e1 && e2, "null evaluator");
        require(npts >= 0 && nbatch >= 0, "negative size");
        if (nbatch == 0) return;
        require(coef && gcoef && work, "null data pointer");
        require(npts == 0 || (coords && w && gx), "null data pointer");
        ddcx::Rules2D const rules = rules2d(e1, e2);
        ddcx::launch_eval2d_vjp_binned(*e1->bs, *e2->bs, rules, coords, w, gx, npts, coef, gcoef, coef_batch_stride,
                                       st1, st2, nbatch, work, work_bytes, static_cast<hipStream_t>(stream));
    });
}

ddcx_status ddcx_evaluator2d_eval_vjp_binned_static(const ddcx_evaluator* e1, const ddcx_evaluator* e2,
                                                    const double* coords, const double* w, double* gx, int64_t npts,
                                                    const double* coef, double* gcoef, int64_t coef_batch_stride,
                                                    int64_t st1, int64_t st2, int64_t nbatch, void* stream)
{
    return guarded([&] {
        require(e1 && e2, "null evaluator");
        require(npts >= 0 && nbatch >= 0, "negative size");
        if (nbatch == 0) return;
        require(coef && gcoef, "null data pointer");
        require(npts == 0 || (coords && w && gx), "null data pointer");
        int64_t const need = ddcx::eval2d_vjp_workspace(*e1->bs, *e2->bs, npts, nbatch);
        if (need == 0)
            throw ddcx::Error(DDCX_ERR_UNSUPPORTED, "binned 2D adjoint: grid or degree out of range");
        if (need > int64_t(DDCX_STATIC_SCRATCH_BYTES))
            throw ddcx::Error(DDCX_ERR_INVALID, ("binned 2D adjoint: the call needs " + std::to_string(need) +
                                                 " bytes of workspace, above the static workspace of " +
                                                 std::to_string(int64_t(DDCX_STATIC_SCRATCH_BYTES)) + " bytes"));
        ddcx::Rules2D const rules = rules2d(e1, e2);
        ddcx::launch_eval2d_vjp_binned(*e1->bs, *e2->bs, rules, coords, w, gx, npts, coef, gcoef, coef_batch_stride,
                                       st1, st2, nbatch, ddcx::static_scratch(), int64_t(DDCX_STATIC_SCRATCH_BYTES),
                                       static_cast<hipStream_t>(stream));
    });
}

namespace {

// the evaluators' plans and rules, checked as ddcx_evaluatornd_eval requires
void nd_setup(const ddcx_evaluator* const* evs, int nd, const int* orders, ddcx::BSplinesPlan const** bs, int* lower,
              int* upper)
{
    require(evs && nd >= 1 && nd <= ddcx::kMaxNd, "nd must be between 1 and 4");
    for (int k = 0; k < nd; ++k) {
        require(evs[k], "null evaluator");
        bs[k] = evs[k]->bs.get();
        lower[k] = rule_of(evs[k]->lower);
        upper[k] = rule_of(evs[k]->upper);
        if (!bs[k]->view.periodic && (lower[k] == DDCX_EXTRAP_CONSTANT || upper[k] == DDCX_EXTRAP_CONSTANT))
            throw ddcx::Error(DDCX_ERR_UNSUPPORTED, "the N-dimensional evaluator takes Null or Periodic extrapolation rules");
        if (bs[k]->view.degree > 5)
            throw ddcx::Error(DDCX_ERR_UNSUPPORTED, "N-dimensional evaluation supports degrees 1..5");
        if (orders) {
            require(orders[k] >= 0, "negative derivative order");
            require(orders[k] <= bs[k]->view.degree, "derivative order above the degree");
        }
    }
}

} // namespace

ddcx_status ddcx_evaluatornd_eval(const ddcx_evaluator* const* evs, int nd, const int* orders, double* out,
                                  const double* coords, int64_t npts, const double* coef, int64_t coef_batch_stride,
                                  const int64_t* coef_strides, int64_t nbatch, void* stream)
{
    return guarded([&] {
        ddcx::BSplinesPlan const* bs[ddcx::kMaxNd];
        int lower[ddcx::kMaxNd], upper[ddcx::kMaxNd];
        nd_setup(evs, nd, orders, bs, lower, upper);
        require(coef_strides, "null coefficient strides");
        require(npts >= 0 && nbatch >= 0, "negative size");
        if (npts * nbatch == 0) return;
        require(out && coords && coef, "null data pointer");
        ddcx::launch_evalnd(bs, nd, orders, lower, upper, out, coords, npts, coef, coef_batch_stride, coef_strides,
                            nbatch, static_cast<hipStream_t>(stream));
    });
}

ddcx_status ddcx_evaluatornd_eval_workspace(const ddcx_evaluator* const* evs, int nd, int64_t npts, int64_t nbatch,
                                            int64_t* bytes)
{
    return guarded([&] {
        ddcx::BSplinesPlan const* bs[ddcx::kMaxNd];
        int lower[ddcx::kMaxNd], upper[ddcx::kMaxNd];
        nd_setup(evs, nd, nullptr, bs, lower, upper);
        require(bytes, "null argument");
        require(npts >= 0 && nbatch >= 0, "negative size");
        *bytes = ddcx::evalnd_binned_workspace(bs, nd, npts, nbatch);
    });
}

ddcx_status ddcx_evaluatornd_eval_binned(const ddcx_evaluator* const* evs, int nd, const int* orders, double* out,
                                         const double* coords, int64_t npts, const double* coef,
                                         int64_t coef_batch_stride, const int64_t* coef_strides, int64_t nbatch,
                                         void* work, int64_t work_bytes, void* stream)
{
    return guarded([&] {
        ddcx::BSplinesPlan const* bs[ddcx::kMaxNd];
        int lower[ddcx::kMaxNd], upper[ddcx::kMaxNd];
        nd_setup(evs, nd, orders, bs, lower, upper);
        require(coef_strides, "null coefficient strides");
        require(npts >= 0 && nbatch >= 0, "negative size");
        int64_t const need = ddcx::evalnd_binned_workspace(bs, nd, npts, nbatch);
        if (need == 0)
            throw ddcx::Error(DDCX_ERR_UNSUPPORTED, "binned N-D evaluation: 3 dimensions of one degree <= 5 only");
        require(work_bytes >= need, "binned N-D evaluation: workspace smaller than ddcx_evaluatornd_eval_workspace");
        require(work && reinterpret_cast<uintptr_t>(work) % 8 == 0, "binned N-D evaluation: workspace null or misaligned");
        if (npts * nbatch == 0) return;
        require(out && coords && coef, "null data pointer");
        ddcx::launch_evalnd_binned(bs, nd, orders, lower, upper, out, coords, npts, coef, coef_batch_stride,
                                   coef_strides, nbatch, work, work_bytes, static_cast<hipStream_t>(stream));
    });
}

ddcx_status ddcx_evaluatornd_integrate(const ddcx_evaluator* const* evs, int nd, double* out, int64_t out_stride,
                                       const double* coef, int64_t coef_batch_stride, const int64_t* coef_strides,
                                       int64_t nbatch, void* stream)
{
    return guarded([&] {
        require(evs && nd >= 1 && nd <= ddcx::kMaxNd, "nd must be between 1 and 4");
        require(coef_strides, "null coefficient strides");
        require(nbatch >= 0, "negative size");
        ddcx::BSplinesPlan const* bs[ddcx::kMaxNd];
        for (int k = 0; k < nd; ++k) {
            require(evs[k], "null evaluator");
            bs[k] = evs[k]->bs.get();
        }
        if (nbatch == 0) return;
        require(out && coef, "null data pointer");
        ddcx::launch_integratend(bs, nd, out, out_stride, coef, coef_batch_stride, coef_strides, nbatch, nullptr, 1,
                                 static_cast<hipStream_t>(stream));
    });
}

ddcx_status ddcx_evaluatornd_integrate_workspace(const ddcx_evaluator* const* evs, int nd, int64_t nbatch,
                                                 int64_t* bytes)
{
    return guarded([&] {
        require(evs && nd >= 1 && nd <= ddcx::kMaxNd, "nd must be between 1 and 4");
        require(bytes, "null argument");
        require(nbatch >= 0, "negative size");
        ddcx::BSplinesPlan const* bs[ddcx::kMaxNd];
        for (int k = 0; k < nd; ++k) {
            require(evs[k], "null evaluator");
            bs[k] = evs[k]->bs.get();
        }
        int const S = ddcx::integratend_split(bs, nd, nbatch);
        *bytes = S > 1 ? nbatch * S * int64_t(sizeof(double)) : 0;
    });
}

ddcx_status ddcx_evaluatornd_integrate_ws(const ddcx_evaluator* const* evs, int nd, double* out, int64_t out_stride,
                                          const double* coef, int64_t coef_batch_stride, const int64_t* coef_strides,
                                          int64_t nbatch, void* work, int64_t work_bytes, void* stream)
{
    return guarded([&] {
        require(evs && nd >= 1 && nd <= ddcx::kMaxNd, "nd must be between 1 and 4");
        require(coef_strides, "null coefficient strides");
        require(nbatch >= 0, "negative size");
        ddcx::BSplinesPlan const* bs[ddcx::kMaxNd];
        for (int k = 0; k < nd; ++k) {
            require(evs[k], "null evaluator");
            bs[k] = evs[k]->bs.get();
        }
        if (nbatch == 0) return;
        require(out && coef, "null data pointer");
        int const S = ddcx::integratend_split(bs, nd, nbatch);
        int64_t const need = S > 1 ? nbatch * S * int64_t(sizeof(double)) : 0;
        require(work_bytes >= need, "integrate workspace smaller than ddcx_evaluatornd_integrate_workspace");
        require(need == 0 || (work && reinterpret_cast<uintptr_t>(work) % alignof(double) == 0),
                "integrate workspace null or misaligned");
        ddcx::launch_integratend(bs, nd, out, out_stride, coef, coef_batch_stride, coef_strides, nbatch,
                                 static_cast<double*>(work), S, static_cast<hipStream_t>(stream));
    });
}

ddcx_status ddcx_problem_create(const ddcx_problem_desc* desc, ddcx_problem** out)
{
    return guarded([&] {
        require(desc && out, "null argument");
        *out = nullptr;
        require(desc->n >= 1, "matrix size must be positive");
        auto h = std::make_unique<ddcx_problem>();
        ddcx::HostProblem& pr = h->p.prob;
        bool const pds = desc->pds != 0;
        switch (desc->kind) {
        case DDCX_PROBLEM_DENSE:
            pr.n = desc->n;
            pr.q.init(ddcx::QType::Dense, desc->n, 0, 0);
            break;
        case DDCX_PROBLEM_BAND:
            // the maker's dense fallback (2 kl + ku + 1 >= n) takes any bandwidths
            // (splines_linear_problem_maker.hpp:56-77); symmetric storage needs kl < n
            require(desc->kl >= 0 && desc->ku >= 0 &&
                        ((desc->kl < desc->n && desc->ku < desc->n) ||
                         (!(pds && desc->kl == desc->ku) && 2 * desc->kl + desc->ku + 1 >= desc->n)),
                    "bad bandwidths");
            pr.n = desc->n;
            ddcx::make_band_q(pr.q, desc->n, desc->kl, desc->ku, pds);
            break;
        case DDCX_PROBLEM_BLOCK_BAND:
            require(desc->kl >= 0 && desc->ku >= 0, "bad bandwidths");
            ddcx::make_block_problem(pr, desc->n, desc->kl, desc->ku, pds, desc->bottom_size, desc->top_size);
            break;
        case DDCX_PROBLEM_PERIODIC_BAND:
            require(desc->kl >= 0 && desc->ku >= 0, "bad bandwidths");
            ddcx::make_periodic_band_problem(pr, desc->n, desc->kl, desc->ku, pds);
            break;
        case DDCX_PROBLEM_BAND_STORAGE:
            require(desc->kl >= 0 && desc->ku >= 0 && desc->kl < desc->n && desc->ku < desc->n, "bad bandwidths");
            pr.n = desc->n;
            pr.q.init(ddcx::QType::Band, desc->n, desc->kl, desc->ku);
            break;
        case DDCX_PROBLEM_BLOCK_DENSE:
            ddcx::make_dense_block_problem(pr, desc->n, desc->bottom_size, desc->top_size);
            break;
        default: throw ddcx::Error(DDCX_ERR_INVALID, "unknown problem kind");
        }
        *out = h.release();
    });
}

ddcx_status ddcx_problem_destroy(ddcx_problem* h)
{
    return guarded([&] { delete h; });
}

ddcx_status ddcx_problem_set_element(ddcx_problem* h, int i, int j, double aij)
{
    return guarded([&] {
        require(h, "null problem");
        require(!h->p.factored, "set_element after setup_solver");
        require(i >= 0 && j >= 0 && i < h->p.prob.n && j < h->p.prob.n, "element index out of range");
        h->p.prob.set(i, j, aij);
    });
}

ddcx_status ddcx_problem_get_element(const ddcx_problem* h, int i, int j, double* aij)
{
    return guarded([&] {
        require(h && aij, "null argument");
        ddcx::HostProblem const& pr = h->p.prob;
        require(i >= 0 && j >= 0 && i < pr.n && j < pr.n, "element index out of range");
        if (!h->p.factored) {
            *aij = pr.get(i, j);
            return;
        }
        // the assembled matrix kept by setup(): main block, gamma, delta and
        // lambda (lambda is not modified by setup) in the 3x3 index space
        ddcx::HostProblem orig;
        orig.n = pr.n;
        orig.top = pr.top;
        orig.blocked = pr.blocked;
        orig.q = pr.q_orig;
        if (pr.blocked) {
            orig.schur = pr.schur;
            orig.schur.q = pr.delta_orig;
            orig.gamma = pr.gamma_orig;
            orig.lambda = pr.lambda;
        }
        *aij = orig.get(i, j);
    });
}

ddcx_status ddcx_problem_setup_solver(ddcx_problem* h)
{
    return guarded([&] {
        require(h, "null problem");
        require(!h->p.factored, "setup_solver called twice");
        h->p.prob.setup();
        h->p.factored = true;
        if (ddcx::device_present()) {
            h->p.buf.upload(h->p.prob, h->p.sv);
            h->p.on_device = true;
        }
    });
}

ddcx_status ddcx_problem_get_info(const ddcx_problem* h, ddcx_problem_info* info)
{
    return guarded([&] {
        require(h && info, "null argument");
        ddcx::HostProblem const& pr = h->p.prob;
        info->size = pr.n;
        info->rhs_rows = pr.rhs_rows();
        info->main_kind = int(pr.q.type);
        info->blocked = pr.blocked ? (pr.top > 0 ? 2 : 1) : 0;
        info->main_size = pr.q.n;
        info->pivoted = 0;
        if (h->p.factored) {
            auto piv = [](std::vector<int> const& ip) {
                for (size_t j = 0; j < ip.size(); ++j)
                    if (ip[j] != int(j)) return 1;
                return 0;
            };
            if (pr.q.type == ddcx::QType::Band || pr.q.type == ddcx::QType::Dense) info->pivoted |= piv(pr.q.ipiv);
            if (pr.blocked) info->pivoted |= piv(pr.schur.ipiv);
        }
    });
}

ddcx_status ddcx_problem_solve(const ddcx_problem* h, double* b, int64_t ldb, int64_t nrhs, int transpose,
                               void* stream)
{
    return guarded([&] {
        require(h, "null problem");
        require(h->p.factored, "solve before setup_solver");
        require(nrhs >= 0 && ldb >= nrhs, "ldb must be at least nrhs");
        if (nrhs == 0) return;
        require(b, "null data pointer");
        if (!h->p.on_device) throw ddcx::Error(DDCX_ERR_HIP, "no HIP device was present at setup_solver");
        ddcx::launch_problem_solve(h->p.sv, b, ldb, nrhs, transpose != 0, static_cast<hipStream_t>(stream));
    });
}

} // extern "C"
