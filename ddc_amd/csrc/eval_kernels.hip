// Batched spline evaluation, derivatives, integrals and the evaluator adjoint.
// One thread per evaluation point (the reference runs one thread per batch
// line with a serial loop over points, spline_evaluator.hpp:280-293).  The
// point -> thread mapping follows the memory layout so that coordinate loads
// and output stores are coalesced in both batch-contiguous and
// interpolation-contiguous layouts.
#include <algorithm>

#include "internal.hpp"

namespace ddcx {

namespace {

__device__ __forceinline__ int64_t lay(Layout const& L, int64_t r, int64_t l)
{
    return r * L.rs + (l % L.inner) * L.is + (l / L.inner) * L.os;
}

// eval_no_bc, spline_evaluator.hpp:556-605
template <int D, int ORDER>
__device__ __forceinline__ double eval_no_bc(BSplinesView const& b, double x, const double* __restrict__ c, int64_t rs)
{
    double v[D + 1];
    int jmin;
    if constexpr (ORDER == 0) {
        jmin = eval_basis<D, true>(b, b.knots, x, v);
    } else {
        double der[(D + 1) * (ORDER + 1)];
        jmin = eval_basis_and_n_derivs<D, ORDER, true>(b, b.knots, x, der);
#pragma unroll
        for (int i = 0; i < D + 1; ++i) v[i] = der[i * (ORDER + 1) + ORDER];
    }
    double y = 0.0;
#pragma unroll
    for (int i = 0; i < D + 1; ++i) y += c[int64_t(jmin + i) * rs] * v[i];
    return y;
}

template <int D, int ORDER>
__global__ void __launch_bounds__(256) eval_kernel(BSplinesView b, Rules1D rules, double* __restrict__ out,
                                                   const double* __restrict__ coords, Layout xl, Layout pl,
                                                   int64_t npts,
                                                   const double* __restrict__ coef, Layout cl, int64_t nbatch,
                                                   int point_fastest)
{
    int64_t const t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= npts * nbatch) return;
    int64_t const i = point_fastest ? t % npts : t / nbatch;
    int64_t const l = point_fastest ? t / npts : t % nbatch;
    int64_t const o = lay(pl, i, l);
    double x = coords[lay(xl, i, l)];
    const double* c = coef + lay(cl, 0, l);
    double y;
    if constexpr (ORDER == 0) {
        bool outside;
        y = apply_bc(b, rules, x, outside) ? eval_no_bc<D, 0>(b, x, c, cl.rs) : 0.0;
    } else {
        y = eval_no_bc<D, ORDER>(b, x, c, cl.rs);
    }
    out[o] = y;
}

// Adjoint of the value form: gx = w * f'(x) inside the domain (the periodic
// wrap has slope 1), 0 where an extrapolation rule applies; the coefficient
// gradient is scattered with atomics (each line's coefficients are a small,
// cache-resident array).
template <int D>
__global__ void __launch_bounds__(256) eval_vjp_kernel(BSplinesView b, Rules1D rules,
                                                       const double* __restrict__ coords, const double* __restrict__ w,
                                                       double* __restrict__ gx, Layout pl, int64_t npts,
                                                       const double* __restrict__ coef, double* __restrict__ gcoef,
                                                       Layout cl, int64_t nbatch, int point_fastest)
{
    int64_t const t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= npts * nbatch) return;
    int64_t const i = point_fastest ? t % npts : t / nbatch;
    int64_t const l = point_fastest ? t / npts : t % nbatch;
    int64_t const o = lay(pl, i, l);
    double x = coords[o];
    double const wi = w[o];
    bool outside;
    if (!apply_bc(b, rules, x, outside)) {
        gx[o] = 0.0;
        return;
    }
    double der[(D + 1) * 2];
    int const jmin = eval_basis_and_n_derivs<D, 1, true>(b, b.knots, x, der);
    const double* c = coef + lay(cl, 0, l);
    double* gc = gcoef + lay(cl, 0, l);
    double dy = 0.0;
#pragma unroll
    for (int k = 0; k < D + 1; ++k) {
        atomicAdd(&gc[int64_t(jmin + k) * cl.rs], wi * der[k * 2]);
        dy += c[int64_t(jmin + k) * cl.rs] * der[k * 2 + 1];
    }
    gx[o] = outside ? 0.0 : wi * dy;
}

__global__ void zero_coef_kernel(double* __restrict__ gcoef, Layout cl, int size, int64_t nbatch)
{
    int64_t const t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= int64_t(size) * nbatch) return;
    int64_t const r = t / nbatch, l = t % nbatch;
    gcoef[lay(cl, r, l)] = 0.0;
}

// SplineEvaluator::integrate, spline_evaluator.hpp:493-526
__global__ void __launch_bounds__(256) integrate_kernel(const double* __restrict__ ib, int size,
                                                        double* __restrict__ out, int64_t ostride,
                                                        const double* __restrict__ coef, Layout cl, int64_t nbatch)
{
    int64_t const l = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (l >= nbatch) return;
    const double* c = coef + lay(cl, 0, l);
    double acc = 0.0;
    for (int i = 0; i < size; ++i) acc += c[int64_t(i) * cl.rs] * ib[i];
    out[l * ostride] = acc;
}

// SplineEvaluator2D::eval / eval_no_bc, spline_evaluator_2d.hpp:679-825.
// Grid: x over the points of one batch entry, y over batch entries.  The
// patch sum is y = sum_i v1[i] * (sum_j c(i, j) v2[j]) with fused
// multiply-adds (a rounding-level reordering of eval_no_bc's
// sum_ij c v1 v2).  UNIT: unit coefficient stride along dimension 2, so each
// patch row is one contiguous run (wide loads).
template <int D1, int D2, int O1, int O2, bool UNIT>
__global__ void __launch_bounds__(256) eval2d_kernel(BSplinesView b1, BSplinesView b2, Rules2D rules,
                                                     double* __restrict__ out, const double* __restrict__ coords,
                                                     int64_t npts, const double* __restrict__ coef, int64_t cbs,
                                                     int64_t st1, int64_t st2, int64_t batch0)
{
    int64_t const p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= npts) return;
    int64_t const l = batch0 + blockIdx.y;
    int64_t const t = l * npts + p;
    double x1 = coords[2 * t], x2 = coords[2 * t + 1];
    if constexpr (O1 == 0 && O2 == 0) {
        bool in1, in2;
        if (!place2d(b1, b2, rules, x1, x2, in1, in2)) {
            out[t] = 0.0;
            return;
        }
    }
    double v1[D1 + 1], v2[D2 + 1];
    int j1, j2;
    if constexpr (O1 == 0) j1 = eval_basis<D1, true>(b1, b1.knots, x1, v1);
    else {
        double der[(D1 + 1) * (O1 + 1)];
        j1 = eval_basis_and_n_derivs<D1, O1, true>(b1, b1.knots, x1, der);
#pragma unroll
        for (int i = 0; i < D1 + 1; ++i) v1[i] = der[i * (O1 + 1) + O1];
    }
    if constexpr (O2 == 0) j2 = eval_basis<D2, true>(b2, b2.knots, x2, v2);
    else {
        double der[(D2 + 1) * (O2 + 1)];
        j2 = eval_basis_and_n_derivs<D2, O2, true>(b2, b2.knots, x2, der);
#pragma unroll
        for (int i = 0; i < D2 + 1; ++i) v2[i] = der[i * (O2 + 1) + O2];
    }
    const double* c = coef + l * cbs + int64_t(j1) * st1 + (UNIT ? int64_t(j2) : int64_t(j2) * st2);
    double y = 0.0;
#pragma unroll
    for (int i = 0; i < D1 + 1; ++i) {
        const double* row = c + int64_t(i) * st1;
        double ri = 0.0;
#pragma unroll
        for (int j = 0; j < D2 + 1; ++j) ri = fma(row[UNIT ? j : int64_t(j) * st2], v2[j], ri);
        y = fma(v1[i], ri, y);
    }
    out[t] = y;
}

// Adjoint of the 2D value form: gx = w * (df/dx1, df/dx2) where the
// coordinate reaches the spline (a coordinate clamped by the constant rule
// has zero gradient, the Null rule zeroes both), and the coefficient
// gradient w * B1_i(x1) B2_j(x2) scattered with atomics.
template <int D1, int D2>
__global__ void __launch_bounds__(256) eval2d_vjp_kernel(BSplinesView b1, BSplinesView b2, Rules2D rules,
                                                         const double* __restrict__ coords,
                                                         const double* __restrict__ w, double* __restrict__ gx,
                                                         int64_t npts, const double* __restrict__ coef,
                                                         double* __restrict__ gcoef, int64_t cbs, int64_t st1,
                                                         int64_t st2, int64_t nbatch)
{
    int64_t const t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= npts * nbatch) return;
    int64_t const l = t / npts;
    double x1 = coords[2 * t];
    double x2 = coords[2 * t + 1];
    double const wt = w[t];
    bool in1, in2;
    if (!place2d(b1, b2, rules, x1, x2, in1, in2)) {
        gx[2 * t] = 0.0;
        gx[2 * t + 1] = 0.0;
        return;
    }
    double e1[(D1 + 1) * 2], e2[(D2 + 1) * 2];
    int const j1 = eval_basis_and_n_derivs<D1, 1, true>(b1, b1.knots, x1, e1);
    int const j2 = eval_basis_and_n_derivs<D2, 1, true>(b2, b2.knots, x2, e2);
    const double* c = coef + l * cbs;
    double* gc = gcoef + l * cbs;
    // f1 = sum_i B1'_i sum_j c_ij B2_j, f2 = sum_i B1_i sum_j c_ij B2'_j (the
    // binned adjoint forms the same sums in the same order)
    double f1 = 0.0, f2 = 0.0;
#pragma unroll
    for (int i = 0; i < D1 + 1; ++i) {
        double s0 = 0.0, s1 = 0.0;
        double const we = wt * e1[i * 2];
#pragma unroll
        for (int j = 0; j < D2 + 1; ++j) {
            int64_t const o = int64_t(j1 + i) * st1 + int64_t(j2 + j) * st2;
            double const cij = c[o];
            s0 = fma(cij, e2[j * 2], s0);
            s1 = fma(cij, e2[j * 2 + 1], s1);
            atomicAdd(&gc[o], we * e2[j * 2]);
        }
        f1 = fma(e1[i * 2 + 1], s0, f1);
        f2 = fma(e1[i * 2], s1, f2);
    }
    gx[2 * t] = in1 ? wt * f1 : 0.0;
    gx[2 * t + 1] = in2 ? wt * f2 : 0.0;
}

__global__ void zero_coef2d_kernel(double* __restrict__ g, int s1, int s2, int64_t cbs, int64_t st1, int64_t st2,
                                   int64_t nbatch)
{
    int64_t const t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    int64_t const per = int64_t(s1) * s2;
    if (t >= per * nbatch) return;
    int64_t const l = t / per, r = t % per;
    g[l * cbs + (r / s2) * st1 + (r % s2) * st2] = 0.0;
}

unsigned grid_for(int64_t n, int threads)
{
    int64_t const b = (n + threads - 1) / threads;
    if (b > int64_t(0x7fffffff)) throw Error(DDCX_ERR_INVALID, "batch too large for one launch");
    return unsigned(b);
}

template <int D, int O>
void eval_launch(BSplinesView const& b, Rules1D rules, double* out, const double* coords, Layout xl, Layout pl,
                 int64_t npts, const double* coef, Layout cl, int64_t nbatch, hipStream_t stream)
{
    int const point_fastest = pl.rs == 1 ? 1 : 0;
    hipLaunchKernelGGL((eval_kernel<D, O>), dim3(grid_for(npts * nbatch, 256)), dim3(256), 0, stream, b, rules,
                       out, coords, xl, pl, npts, coef, cl, nbatch, point_fastest);
}

template <int D>
void eval_order(int order, BSplinesView const& b, Rules1D rules, double* out, const double* coords, Layout xl,
                Layout pl, int64_t npts, const double* coef, Layout cl, int64_t nbatch, hipStream_t stream)
{
    switch (order) {
    case 0: eval_launch<D, 0>(b, rules, out, coords, xl, pl, npts, coef, cl, nbatch, stream); return;
    case 1: eval_launch<D, 1>(b, rules, out, coords, xl, pl, npts, coef, cl, nbatch, stream); return;
    case 2: if constexpr (D >= 2) { eval_launch<D, 2>(b, rules, out, coords, xl, pl, npts, coef, cl, nbatch, stream); return; } break;
    case 3: if constexpr (D >= 3) { eval_launch<D, 3>(b, rules, out, coords, xl, pl, npts, coef, cl, nbatch, stream); return; } break;
    case 4: if constexpr (D >= 4) { eval_launch<D, 4>(b, rules, out, coords, xl, pl, npts, coef, cl, nbatch, stream); return; } break;
    case 5: if constexpr (D >= 5) { eval_launch<D, 5>(b, rules, out, coords, xl, pl, npts, coef, cl, nbatch, stream); return; } break;
    case 6: if constexpr (D >= 6) { eval_launch<D, 6>(b, rules, out, coords, xl, pl, npts, coef, cl, nbatch, stream); return; } break;
    case 7: if constexpr (D >= 7) { eval_launch<D, 7>(b, rules, out, coords, xl, pl, npts, coef, cl, nbatch, stream); return; } break;
    default: break;
    }
    throw Error(DDCX_ERR_INVALID, "derivative order must be in [1, degree]");
}

} // namespace

void launch_eval(BSplinesPlan const& bs, Rules1D rules, int order, double* out, const double* coords,
                 Layout pl, int64_t npts, const double* coef, Layout cl, int64_t nbatch, hipStream_t stream)
{
    launch_eval_xl(bs, rules, order, out, coords, pl, pl, npts, coef, cl, nbatch, stream);
}

void launch_eval_xl(BSplinesPlan const& bs, Rules1D rules, int order, double* out, const double* coords,
                    Layout xl, Layout pl, int64_t npts, const double* coef, Layout cl, int64_t nbatch,
                    hipStream_t stream)
{
    if (npts * nbatch == 0) return;
    bs.ensure_device();
    if (launch_eval_tiled(bs, rules, order, out, coords, xl, pl, npts, coef, cl, nbatch, stream)) return;
    BSplinesView const b = bs.device_view();
    bool const ok = dispatch_degree(b.degree, [&]<int D>() {
        eval_order<D>(order, b, rules, out, coords, xl, pl, npts, coef, cl, nbatch, stream);
    });
    if (!ok) throw Error(DDCX_ERR_UNSUPPORTED, "degree out of range");
    hip_check(hipGetLastError(), "eval_kernel launch");
}

void launch_eval_vjp(BSplinesPlan const& bs, Rules1D rules, const double* coords, const double* w, double* gx,
                     Layout pl, int64_t npts, const double* coef, double* gcoef, Layout cl, int64_t nbatch,
                     hipStream_t stream)
{
    bs.ensure_device();
    if (launch_eval_vjp_tiled(bs, rules, coords, w, gx, pl, npts, coef, gcoef, cl, nbatch, stream)) return;
    BSplinesView const b = bs.device_view();
    int const size = bs.size();
    if (int64_t(size) * nbatch > 0)
        hipLaunchKernelGGL(zero_coef_kernel, dim3(grid_for(int64_t(size) * nbatch, 256)), dim3(256), 0, stream, gcoef,
                           cl, size, nbatch);
    if (npts * nbatch == 0) return;
    int const point_fastest = pl.rs == 1 ? 1 : 0;
    bool const ok = dispatch_degree(b.degree, [&]<int D>() {
        hipLaunchKernelGGL(eval_vjp_kernel<D>, dim3(grid_for(npts * nbatch, 256)), dim3(256), 0, stream, b, rules,
                           coords, w, gx, pl, npts, coef, gcoef, cl, nbatch, point_fastest);
    });
    if (!ok) throw Error(DDCX_ERR_UNSUPPORTED, "degree out of range");
    hip_check(hipGetLastError(), "eval_vjp_kernel launch");
}

void launch_integrate(BSplinesPlan const& bs, double* out, int64_t ostride, const double* coef, Layout cl,
                      int64_t nbatch, hipStream_t stream)
{
    if (nbatch == 0) return;
    bs.ensure_device();
    hipLaunchKernelGGL(integrate_kernel, dim3(grid_for(nbatch, 256)), dim3(256), 0, stream,
                       bs.integrals_dev.as<double>(), bs.size(), out, ostride, coef, cl, nbatch);
    hip_check(hipGetLastError(), "integrate_kernel launch");
}

void launch_eval2d(BSplinesPlan const& bp1, BSplinesPlan const& bp2, Rules2D const& rules, int d1, int d2, double* out,
                   const double* coords, int64_t npts, const double* coef, int64_t cbs, int64_t st1, int64_t st2,
                   int64_t nbatch, hipStream_t stream)
{
    if (npts * nbatch == 0) return;
    bp1.ensure_device();
    bp2.ensure_device();
    BSplinesView const b1 = bp1.device_view(), b2 = bp2.device_view();
    unsigned const gx = grid_for(npts, 256);
    bool const unit = st2 == 1;
    bool launched = false;
    auto go = [&]<int D1, int D2>() {
        auto ord = [&]<int O1, int O2>() {
            for (int64_t l0 = 0; l0 < nbatch; l0 += 65535) {
                unsigned const gy = unsigned(std::min<int64_t>(65535, nbatch - l0));
                if (unit)
                    hipLaunchKernelGGL((eval2d_kernel<D1, D2, O1, O2, true>), dim3(gx, gy), dim3(256), 0, stream, b1,
                                       b2, rules, out, coords, npts, coef, cbs, st1, st2, l0);
                else
                    hipLaunchKernelGGL((eval2d_kernel<D1, D2, O1, O2, false>), dim3(gx, gy), dim3(256), 0, stream,
                                       b1, b2, rules, out, coords, npts, coef, cbs, st1, st2, l0);
            }
            launched = true;
        };
        auto sel = [&]<int O1, int O2>() {
            if constexpr (O1 <= D1 && O2 <= D2) {
                if (!launched && d1 == O1 && d2 == O2) ord.template operator()<O1, O2>();
            }
        };
        sel.template operator()<0, 0>();
        sel.template operator()<1, 0>();
        sel.template operator()<0, 1>();
        sel.template operator()<1, 1>();
        sel.template operator()<2, 0>();
        sel.template operator()<0, 2>();
        sel.template operator()<2, 1>();
        sel.template operator()<1, 2>();
        sel.template operator()<2, 2>();
    };
    // 2D evaluation is instantiated for degrees 1..5 in each dimension
    auto outer = [&]<int D1>() {
        switch (b2.degree) {
        case 1: go.template operator()<D1, 1>(); break;
        case 2: go.template operator()<D1, 2>(); break;
        case 3: go.template operator()<D1, 3>(); break;
        case 4: go.template operator()<D1, 4>(); break;
        case 5: go.template operator()<D1, 5>(); break;
        default: break;
        }
    };
    switch (b1.degree) {
    case 1: outer.template operator()<1>(); break;
    case 2: outer.template operator()<2>(); break;
    case 3: outer.template operator()<3>(); break;
    case 4: outer.template operator()<4>(); break;
    case 5: outer.template operator()<5>(); break;
    default: break;
    }
    if (!launched) throw Error(DDCX_ERR_UNSUPPORTED, "2D evaluation supports degrees 1..5 and derivative orders <= 2");
    hip_check(hipGetLastError(), "eval2d_kernel launch");
}

void launch_eval2d_vjp(BSplinesPlan const& bp1, BSplinesPlan const& bp2, Rules2D const& rules, const double* coords,
                       const double* w, double* gx, int64_t npts, const double* coef, double* gcoef, int64_t cbs,
                       int64_t st1, int64_t st2, int64_t nbatch, hipStream_t stream)
{
    if (nbatch == 0) return;
    bp1.ensure_device();
    bp2.ensure_device();
    BSplinesView const b1 = bp1.device_view(), b2 = bp2.device_view();
    hipLaunchKernelGGL(zero_coef2d_kernel, dim3(grid_for(int64_t(bp1.size()) * bp2.size() * nbatch, 256)), dim3(256),
                       0, stream, gcoef, bp1.size(), bp2.size(), cbs, st1, st2, nbatch);
    hip_check(hipGetLastError(), "zero_coef2d_kernel launch");
    if (npts == 0) return;
    unsigned const grid = grid_for(npts * nbatch, 256);
    bool launched = false;
    auto go = [&]<int D1, int D2>() {
        hipLaunchKernelGGL((eval2d_vjp_kernel<D1, D2>), dim3(grid), dim3(256), 0, stream, b1, b2, rules, coords, w, gx,
                           npts, coef, gcoef, cbs, st1, st2, nbatch);
        launched = true;
    };
    auto outer = [&]<int D1>() {
        switch (b2.degree) {
        case 1: go.template operator()<D1, 1>(); break;
        case 2: go.template operator()<D1, 2>(); break;
        case 3: go.template operator()<D1, 3>(); break;
        case 4: go.template operator()<D1, 4>(); break;
        case 5: go.template operator()<D1, 5>(); break;
        default: break;
        }
    };
    switch (b1.degree) {
    case 1: outer.template operator()<1>(); break;
    case 2: outer.template operator()<2>(); break;
    case 3: outer.template operator()<3>(); break;
    case 4: outer.template operator()<4>(); break;
    case 5: outer.template operator()<5>(); break;
    default: break;
    }
    if (!launched) throw Error(DDCX_ERR_UNSUPPORTED, "2D adjoint supports degrees 1..5");
    hip_check(hipGetLastError(), "eval2d_vjp_kernel launch");
}

} // namespace ddcx
