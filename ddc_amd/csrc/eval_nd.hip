// Batched N-dimensional tensor-product evaluation and integration
// (SplineEvaluatorND, spline_evaluator_nd.hpp; SplineEvaluator3D,
// spline_evaluator_3d.hpp, is its N = 3 case).
//
// One thread per point: the per-dimension basis values (degree and
// derivative order are kernel arguments, dispatched to the compile-time
// basis routines by a wave-uniform switch), then the sum over the
// (d_1+1) x ... x (d_N+1) coefficient patch in row-major order with the
// basis product folded from the right, as eval_no_bc does
// (spline_evaluator_nd.hpp:782-826).  Integration reduces
// coef(i) * (I_1(i_1) * (...)) over the coefficient block of each batch entry
// (spline_evaluator_nd.hpp:597-632): one workgroup per entry, or, with a
// workspace for the partial sums, S workgroups per entry and a second pass.
#include "internal.hpp"

#include <algorithm>
#include <utility>

namespace ddcx {

namespace {

struct NdArgs {
    BSplinesView b[kMaxNd];
    const double* integrals[kMaxNd];
    int64_t st[kMaxNd];
    int order[kMaxNd];
    int lower[kMaxNd];
    int upper[kMaxNd];
};

constexpr int kM = 6; // basis values per dimension (degree <= 5)

// y += coef(jmin + idx) * (v_1[idx_1] * (v_2[idx_2] * (...))) over the
// multi-index in row-major order, every index a compile-time constant (so
// the basis tables stay in registers) behind a wave-uniform bound check.
template <int N, int... I>
__device__ __forceinline__ void tensor_acc(double const (&v)[N][kM], int const (&n)[N], int64_t const (&st)[N],
                                           const double* c, double& y)
{
    constexpr int K = sizeof...(I);
    if constexpr (K == N) {
        constexpr int ix[N] = {I...};
        double p = v[N - 1][ix[N - 1]];
        int64_t off = ix[N - 1] * st[N - 1];
#pragma unroll
        for (int k = N - 2; k >= 0; --k) {
            p = v[k][ix[k]] * p;
            off += ix[k] * st[k];
        }
        y += c[off] * p;
    } else {
        [&]<int... J>(std::integer_sequence<int, J...>) {
            ((J < n[K] ? tensor_acc<N, I..., J>(v, n, st, c, y) : void()), ...);
        }(std::make_integer_sequence<int, kM>{});
    }
}

template <int D>
__device__ int basis_of_order(BSplinesView const& b, double x, int order, double* v)
{
    if (order == 0) return eval_basis<D, true>(b, b.knots, x, v);
    int j = 0;
    auto take = [&]<int O>() {
        if constexpr (O <= D) {
            double der[(D + 1) * (O + 1)];
            j = eval_basis_and_n_derivs<D, O, true>(b, b.knots, x, der);
#pragma unroll
            for (int i = 0; i < D + 1; ++i) v[i] = der[i * (O + 1) + O];
        }
    };
    switch (order) {
    case 1: take.template operator()<1>(); break;
    case 2: take.template operator()<2>(); break;
    case 3: take.template operator()<3>(); break;
    case 4: take.template operator()<4>(); break;
    case 5: take.template operator()<5>(); break;
    default: break;
    }
    return j;
}

__device__ int basis_any(BSplinesView const& b, double x, int order, double* v)
{
    switch (b.degree) {
    case 1: return basis_of_order<1>(b, x, order, v);
    case 2: return basis_of_order<2>(b, x, order, v);
    case 3: return basis_of_order<3>(b, x, order, v);
    case 4: return basis_of_order<4>(b, x, order, v);
    case 5: return basis_of_order<5>(b, x, order, v);
    default: return 0;
    }
}

// SplineEvaluatorND::eval (spline_evaluator_nd.hpp:684-720) when every order
// is zero: periodic dimensions wrap first, then the first non-periodic
// dimension outside [rmin, rmax] applies its rule (only Null reaches here:
// 0).  Nonzero orders go straight to eval_no_bc, as deriv() does (:445-456).
template <int N>
__global__ void __launch_bounds__(256) evalnd_kernel(NdArgs a, double* __restrict__ out,
                                                     const double* __restrict__ coords, int64_t npts,
                                                     const double* __restrict__ coef, int64_t cbs, int64_t nbatch,
                                                     int any_order)
{
    int64_t const t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= npts * nbatch) return;
    int64_t const l = t / npts;
    double x[N];
#pragma unroll
    for (int k = 0; k < N; ++k) x[k] = coords[N * t + k];
    if (!any_order) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            BSplinesView const& b = a.b[k];
            if (b.periodic && (x[k] < b.rmin || x[k] > b.rmax)) x[k] -= floor((x[k] - b.rmin) / b.length) * b.length;
        }
#pragma unroll
        for (int k = 0; k < N; ++k) {
            BSplinesView const& b = a.b[k];
            if (!b.periodic && ((x[k] < b.rmin && a.lower[k] == DDCX_EXTRAP_NULL) ||
                                (x[k] > b.rmax && a.upper[k] == DDCX_EXTRAP_NULL))) {
                out[t] = 0.0;
                return;
            }
        }
    }
    double v[N][kM];
    int64_t st[N];
    int64_t base = 0;
    int n[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        int const j = basis_any(a.b[k], x[k], a.order[k], v[k]);
        st[k] = a.st[k];
        base += int64_t(j) * st[k];
        n[k] = a.b[k].degree + 1;
    }
    double y = 0.0;
    tensor_acc<N>(v, n, st, coef + l * cbs + base, y);
    out[t] = y;
}

// Same degree D in every dimension: the basis routines inline with the
// degree fixed, the tables stay in registers and the patch loop has exactly
// (D+1)^N terms, in the same row-major order and right-folded products as
// tensor_acc, so the values are bit-identical to evalnd_kernel's.
template <int N, int D, int... I>
__device__ __forceinline__ void tensor_acc_fixed(double const (&v)[N][kM], int64_t const (&st)[N], const double* c,
                                                 double& y)
{
    constexpr int K = sizeof...(I);
    if constexpr (K == N) {
        constexpr int ix[N] = {I...};
        double p = v[N - 1][ix[N - 1]];
        int64_t off = ix[N - 1] * st[N - 1];
#pragma unroll
        for (int k = N - 2; k >= 0; --k) {
            p = v[k][ix[k]] * p;
            off += ix[k] * st[k];
        }
        y += c[off] * p;
    } else {
        [&]<int... J>(std::integer_sequence<int, J...>) {
            (tensor_acc_fixed<N, D, I..., J>(v, st, c, y), ...);
        }(std::make_integer_sequence<int, D + 1>{});
    }
}

template <int D, bool ANY>
__device__ __forceinline__ int basis_fixed(BSplinesView const& b, double x, int order, double* v)
{
    if constexpr (ANY) {
        if (order != 0) {
            int j = 0;
            auto take = [&]<int O>() {
                if constexpr (O <= D) {
                    double der[(D + 1) * (O + 1)];
                    j = eval_basis_and_n_derivs<D, O, true>(b, b.knots, x, der);
#pragma unroll
                    for (int i = 0; i < D + 1; ++i) v[i] = der[i * (O + 1) + O];
                }
            };
            switch (order) {
            case 1: take.template operator()<1>(); break;
            case 2: take.template operator()<2>(); break;
            case 3: take.template operator()<3>(); break;
            case 4: take.template operator()<4>(); break;
            case 5: take.template operator()<5>(); break;
            default: break;
            }
            return j;
        }
    }
    return eval_basis<D, true>(b, b.knots, x, v);
}

// evalnd_kernel with every dimension of degree D; ANY: some order is nonzero
// (no extrapolation rules then, as in evalnd_kernel)
template <int N, int D, bool ANY>
__global__ void __launch_bounds__(256) evalnd_fixed_kernel(NdArgs a, double* __restrict__ out,
                                                           const double* __restrict__ coords, int64_t npts,
                                                           const double* __restrict__ coef, int64_t cbs,
                                                           int64_t nbatch)
{
    int64_t const t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= npts * nbatch) return;
    int64_t const l = t / npts;
    double x[N];
#pragma unroll
    for (int k = 0; k < N; ++k) x[k] = coords[N * t + k];
    if constexpr (!ANY) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            BSplinesView const& b = a.b[k];
            if (b.periodic && (x[k] < b.rmin || x[k] > b.rmax)) x[k] -= floor((x[k] - b.rmin) / b.length) * b.length;
        }
#pragma unroll
        for (int k = 0; k < N; ++k) {
            BSplinesView const& b = a.b[k];
            if (!b.periodic && ((x[k] < b.rmin && a.lower[k] == DDCX_EXTRAP_NULL) ||
                                (x[k] > b.rmax && a.upper[k] == DDCX_EXTRAP_NULL))) {
                out[t] = 0.0;
                return;
            }
        }
    }
    double v[N][kM];
    int64_t st[N];
    int64_t base = 0;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        int const j = basis_fixed<D, ANY>(a.b[k], x[k], ANY ? a.order[k] : 0, v[k]);
        st[k] = a.st[k];
        base += int64_t(j) * st[k];
    }
    double y = 0.0;
    tensor_acc_fixed<N, D>(v, st, coef + l * cbs + base, y);
    out[t] = y;
}

constexpr int kIntThreads = 1024;

// Workgroup g takes batch entry l = g / S and the rows (all indices but the
// last) [rows*s/S, rows*(s+1)/S) of it, s = g % S.  Wave w takes those rows
// in groups of RW (then the next group of the next round), and its lanes
// sweep the last dimension, so the loads of a row are contiguous when the
// last dimension is.  Each term is coef(i) * (I_1(i_1) * (... * I_N(i_N))) as
// the reference forms it; the per-lane sums are combined by a fixed LDS tree
// into dst[l*dst_ls + s], and the S partial sums of an entry by
// integratend_sum_kernel in index order, so results do not depend on
// scheduling (they do depend on S, which is a function of the sizes only).
template <int N>
__global__ void __launch_bounds__(kIntThreads) integratend_kernel(NdArgs a, double* __restrict__ dst, int64_t dst_ls,
                                                                  const double* __restrict__ coef, int64_t cbs,
                                                                  int S)
{
    constexpr int W = 64, NW = kIntThreads / W;
    __shared__ double part[kIntThreads];
    int64_t const l = blockIdx.x / unsigned(S);
    int const s = int(blockIdx.x - l * S);
    int const lane = threadIdx.x % W;
    int const wave = threadIdx.x / W;
    int n[N];
    int64_t rows = 1;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        n[k] = a.b[k].ncells + a.b[k].degree;
        if (k < N - 1) rows *= n[k];
    }
    int64_t const r_lo = rows * s / S, r_hi = rows * (s + 1) / S;
    const double* c = coef + l * cbs;
    const double* qn = a.integrals[N - 1];
    int64_t const sn = a.st[N - 1];
    // RW rows per wave at a time, with independent loads and sums
    constexpr int RW = 4;
    double acc[RW];
#pragma unroll
    for (int u = 0; u < RW; ++u) acc[u] = 0.0;
    for (int64_t r0 = r_lo + int64_t(wave) * RW; r0 < r_hi; r0 += int64_t(NW) * RW) {
        int64_t off[RW];
        double outer[RW][N > 1 ? N - 1 : 1];
#pragma unroll
        for (int u = 0; u < RW; ++u) {
            uint32_t rr = uint32_t(min(r0 + u, r_hi - 1)); // rows < 2^31, checked at launch
            off[u] = 0;
#pragma unroll
            for (int k = N - 2; k >= 0; --k) {
                int const ik = int(rr % uint32_t(n[k]));
                rr /= uint32_t(n[k]);
                off[u] += ik * a.st[k];
                outer[u][k] = a.integrals[k][ik];
            }
        }
        for (int i = lane; i < n[N - 1]; i += W) {
            double const q = qn[i];
#pragma unroll
            for (int u = 0; u < RW; ++u) {
                double p = q;
#pragma unroll
                for (int k = N - 2; k >= 0; --k) p = outer[u][k] * p;
                double const t = c[off[u] + i * sn] * p;
                if (r0 + u < r_hi) acc[u] += t;
            }
        }
    }
    double sum = 0.0;
#pragma unroll
    for (int u = 0; u < RW; ++u) sum += acc[u];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int h = kIntThreads / 2; h > 0; h /= 2) {
        if (int(threadIdx.x) < h) part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) dst[l * dst_ls + s] = part[0];
}

// out[l*ostride] = sum of the S partial sums of entry l: one wave per entry,
// lane i adds partials i, i + 64, ... in order, then a fixed butterfly
// combines the lanes, so the order of the additions is a function of S only
__global__ void __launch_bounds__(256) integratend_sum_kernel(const double* __restrict__ part, int S,
                                                              double* __restrict__ out, int64_t ostride,
                                                              int64_t nbatch)
{
    int64_t const l = int64_t(blockIdx.x) * 4 + threadIdx.x / 64;
    int const lane = threadIdx.x % 64;
    if (l >= nbatch) return; // wave-uniform
    double y = 0.0;
    for (int s = lane; s < S; s += 64) y += part[l * S + s];
#pragma unroll
    for (int m = 32; m >= 1; m /= 2) y += __shfl_xor(y, m, 64);
    if (lane == 0) out[l * ostride] = y;
}

NdArgs make_args(BSplinesPlan const* const* bs, int nd, const int* orders, const int* lower, const int* upper,
                 const int64_t* st)
{
    NdArgs a{};
    for (int k = 0; k < nd; ++k) {
        bs[k]->ensure_device();
        a.b[k] = bs[k]->device_view();
        a.integrals[k] = bs[k]->integrals_dev.as<double>();
        a.st[k] = st[k];
        a.order[k] = orders ? orders[k] : 0;
        a.lower[k] = lower ? lower[k] : DDCX_EXTRAP_NULL;
        a.upper[k] = upper ? upper[k] : DDCX_EXTRAP_NULL;
    }
    return a;
}

template <class F>
void dispatch_nd(int nd, F&& f)
{
    switch (nd) {
    case 1: f.template operator()<1>(); break;
    case 2: f.template operator()<2>(); break;
    case 3: f.template operator()<3>(); break;
    case 4: f.template operator()<4>(); break;
    default: throw Error(DDCX_ERR_UNSUPPORTED, "N-dimensional evaluation supports 1 to 4 dimensions");
    }
}

} // namespace

void launch_evalnd(BSplinesPlan const* const* bs, int nd, const int* orders, const int* lower, const int* upper,
                   double* out, const double* coords, int64_t npts, const double* coef, int64_t cbs,
                   const int64_t* st, int64_t nbatch, hipStream_t stream)
{
    if (npts * nbatch == 0) return;
    NdArgs const a = make_args(bs, nd, orders, lower, upper, st);
    int any = 0;
    for (int k = 0; k < nd; ++k) any |= a.order[k];
    int64_t const total = npts * nbatch;
    int64_t const blocks = (total + 255) / 256;
    if (blocks > int64_t(0x7fffffff)) throw Error(DDCX_ERR_INVALID, "batch too large for one launch");
    // one degree in every dimension (N = 2, 3, and 4 up to cubic): the fixed-degree kernel
    int deg = a.b[0].degree;
    for (int k = 1; k < nd; ++k)
        if (a.b[k].degree != deg) deg = -1;
    bool launched = false;
    auto fixed = [&]<int N>() {
        auto go = [&]<int D>() {
            if (deg != D) return;
            if (any)
                hipLaunchKernelGGL((evalnd_fixed_kernel<N, D, true>), dim3(unsigned(blocks)), dim3(256), 0, stream, a,
                                   out, coords, npts, coef, cbs, nbatch);
            else
                hipLaunchKernelGGL((evalnd_fixed_kernel<N, D, false>), dim3(unsigned(blocks)), dim3(256), 0, stream,
                                   a, out, coords, npts, coef, cbs, nbatch);
            launched = true;
        };
        go.template operator()<1>();
        go.template operator()<2>();
        go.template operator()<3>();
        if constexpr (N < 4) {
            go.template operator()<4>();
            go.template operator()<5>();
        }
    };
    if (nd == 2) fixed.template operator()<2>();
    else if (nd == 3) fixed.template operator()<3>();
    else if (nd == 4) fixed.template operator()<4>();
    if (launched) {
        hip_check(hipGetLastError(), "evalnd_fixed_kernel launch");
        return;
    }
    dispatch_nd(nd, [&]<int N>() {
        hipLaunchKernelGGL((evalnd_kernel<N>), dim3(unsigned(blocks)), dim3(256), 0, stream, a, out, coords, npts,
                           coef, cbs, nbatch, any);
    });
    hip_check(hipGetLastError(), "evalnd_kernel launch");
}

namespace {

int64_t integratend_rows(BSplinesPlan const* const* bs, int nd)
{
    int64_t rows = 1;
    for (int k = 0; k + 1 < nd; ++k) rows *= bs[k]->size();
    return rows;
}

} // namespace

int integratend_split(BSplinesPlan const* const* bs, int nd, int64_t nbatch)
{
    // workgroups per batch entry: about 8 per CU over the whole batch, each
    // with at least 64 rows
    int64_t const rows = integratend_rows(bs, nd);
    int64_t s = (2048 + std::max<int64_t>(nbatch, 1) - 1) / std::max<int64_t>(nbatch, 1);
    s = std::min<int64_t>(s, std::max<int64_t>(1, rows / 64));
    return int(std::max<int64_t>(1, std::min<int64_t>(s, 1024)));
}

void launch_integratend(BSplinesPlan const* const* bs, int nd, double* out, int64_t ostride, const double* coef,
                        int64_t cbs, const int64_t* st, int64_t nbatch, double* part, int S, hipStream_t stream)
{
    if (nbatch == 0) return;
    NdArgs const a = make_args(bs, nd, nullptr, nullptr, nullptr, st);
    int64_t const rows = integratend_rows(bs, nd);
    if (rows >= (int64_t(1) << 31)) throw Error(DDCX_ERR_UNSUPPORTED, "too many coefficient rows to integrate");
    if (S < 1 || (S > 1 && !part)) throw Error(DDCX_ERR_INVALID, "integrate: bad split");
    if (nbatch * S > int64_t(0x7fffffff)) throw Error(DDCX_ERR_INVALID, "batch too large for one launch");
    dispatch_nd(nd, [&]<int N>() {
        if (S == 1)
            hipLaunchKernelGGL((integratend_kernel<N>), dim3(unsigned(nbatch)), dim3(kIntThreads), 0, stream, a, out,
                               ostride, coef, cbs, 1);
        else
            hipLaunchKernelGGL((integratend_kernel<N>), dim3(unsigned(nbatch * S)), dim3(kIntThreads), 0, stream, a,
                               part, int64_t(S), coef, cbs, S);
    });
    hip_check(hipGetLastError(), "integratend_kernel launch");
    if (S > 1) {
        hipLaunchKernelGGL(integratend_sum_kernel, dim3(unsigned((nbatch + 3) / 4)), dim3(256), 0, stream, part, S,
                           out, ostride, nbatch);
        hip_check(hipGetLastError(), "integratend_sum_kernel launch");
    }
}

} // namespace ddcx
