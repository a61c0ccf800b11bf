// ddcx_device.hpp -- device-callable single-point evaluation, for use inside
// the caller's own HIP kernels: SplineEvaluator's operator()(coord, coef_1d)
// and deriv(order, coord, coef_1d) (spline_evaluator.hpp:228-235, :357-478).
// The evaluator's device data come from ddcx_evaluator_device_view (ddcx.h)
// and stay valid while the evaluator exists; coef_1d is one line's
// coefficients, element j at coef[j * stride].  The basis routines are the
// library's own (ddc_amd/csrc/basis.hpp), so the values equal the batched
// evaluator's: bit for bit when this code is compiled, as the library is,
// with -ffp-contract=off, else to rounding.
//
//   __global__ void k(ddcx_device_evaluator de, const double* coef, ...)
//   {
//       ddcx::device::Evaluator const e = ddcx::device::evaluator(de);
//       double const f = ddcx::device::eval(e, x, coef + line, nlines);
//       double const df = ddcx::device::deriv(e, 1, x, coef + line, nlines);
//   }
#pragma once

#include <hip/hip_runtime.h>

#include "../ddc_amd/csrc/basis.hpp"
#include "ddcx.h"

namespace ddcx {
namespace device {

struct Evaluator {
    BSplinesView b;
    Rules1D rules;
};

__host__ __device__ inline Evaluator evaluator(ddcx_device_evaluator const& d)
{
    Evaluator e{};
    e.b.degree = d.degree;
    e.b.uniform = d.uniform;
    e.b.periodic = d.periodic;
    e.b.ncells = d.ncells;
    e.b.origin = d.origin;
    e.b.step = d.step;
    e.b.inv_step = d.inv_step;
    e.b.rmin = d.rmin;
    e.b.rmax = d.rmax;
    e.b.length = d.length;
    e.b.knots = d.knots;
    e.b.cells = d.cells;
    e.b.nbucket = d.nbucket;
    e.b.inv_bucket = d.inv_bucket;
    e.rules = Rules1D{d.lower, d.upper, d.pos_lower, d.pos_upper};
    return e;
}

// order 0: the rules first (periodic wrap, Null: 0, Constant: the rule's
// position), as operator() does; order > 0: no rule, as deriv() does
template <int D, int ORDER>
__device__ inline double eval_fixed(Evaluator const& e, double x, const double* coef, int64_t stride)
{
    double v[D + 1];
    int jmin;
    if constexpr (ORDER == 0) {
        bool outside;
        if (!apply_bc(e.b, e.rules, x, outside)) return 0.0;
        jmin = eval_basis<D, true>(e.b, e.b.knots, x, v);
    } else {
        double der[(D + 1) * (ORDER + 1)];
        jmin = eval_basis_and_n_derivs<D, ORDER, true>(e.b, e.b.knots, x, der);
#pragma unroll
        for (int i = 0; i < D + 1; ++i) v[i] = der[i * (ORDER + 1) + ORDER];
    }
    double y = 0.0;
#pragma unroll
    for (int i = 0; i < D + 1; ++i) y += coef[int64_t(jmin + i) * stride] * v[i];
    return y;
}

// deriv(order, coord, coef_1d); order 0 is operator()(coord, coef_1d).
// Degrees 1..7, orders 0..degree; anything else gives NaN.
__device__ inline double deriv(Evaluator const& e, int order, double x, const double* coef, int64_t stride = 1)
{
    double y = __builtin_nan("");
    auto orders = [&]<int D>() {
        auto one = [&]<int O>() {
            if constexpr (O <= D)
                if (order == O) y = eval_fixed<D, O>(e, x, coef, stride);
        };
        one.template operator()<0>();
        one.template operator()<1>();
        one.template operator()<2>();
        one.template operator()<3>();
        one.template operator()<4>();
        one.template operator()<5>();
        one.template operator()<6>();
        one.template operator()<7>();
    };
    switch (e.b.degree) {
    case 1: orders.template operator()<1>(); break;
    case 2: orders.template operator()<2>(); break;
    case 3: orders.template operator()<3>(); break;
    case 4: orders.template operator()<4>(); break;
    case 5: orders.template operator()<5>(); break;
    case 6: orders.template operator()<6>(); break;
    case 7: orders.template operator()<7>(); break;
    default: break;
    }
    return y;
}

__device__ inline double eval(Evaluator const& e, double x, const double* coef, int64_t stride = 1)
{
    return deriv(e, 0, x, coef, stride);
}

} // namespace device
} // namespace ddcx
