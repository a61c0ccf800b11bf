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
#include "binning.hpp"

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

// sum over the multi-index of coef(jmin + idx) * v_1[idx_1] * ... * v_N[idx_N],
// contracted one dimension at a time from the last: for each index of the
// leading dimensions, s = c_0 v_0, then s = fma(c_i, v_i, s) for i = 1..n-1,
// and the same over the next dimension up with those sums as coefficients.
// A tricubic point takes 84 multiply-adds in chains of 4 instead of 208
// operations behind one 64-term addition chain. Every index is a
// compile-time constant (the basis tables stay in registers); the bounds
// n[K] are wave-uniform.
template <int N, int K = 0>
__device__ __forceinline__ double tensor_contract(double const (&v)[N][kM], int const (&n)[N], int64_t const (&st)[N],
                                                  const double* c)
{
    auto term = [&](int i) {
        if constexpr (K == N - 1) return c[i * st[K]];
        else return tensor_contract<N, K + 1>(v, n, st, c + i * st[K]);
    };
    double s = term(0) * v[K][0];
#pragma unroll
    for (int i = 1; i < kM; ++i)
        if (i < n[K]) s = fma(term(i), v[K][i], s);
    return s;
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
    out[t] = tensor_contract<N>(v, n, st, coef + l * cbs + base);
}

// Same degree D in every dimension: the basis routines inline with the
// degree fixed and the contraction has exactly D+1 terms per dimension, in
// the order and with the operations of tensor_contract, so the values are
// bit-identical to evalnd_kernel's.
template <int N, int D, int K = 0>
__device__ __forceinline__ double tensor_contract_fixed(double const (&v)[N][kM], int64_t const (&st)[N],
                                                        const double* c)
{
    auto term = [&](int i) {
        if constexpr (K == N - 1) return c[i * st[K]];
        else return tensor_contract_fixed<N, D, K + 1>(v, st, c + i * st[K]);
    };
    double s = term(0) * v[K][0];
#pragma unroll
    for (int i = 1; i <= D; ++i) s = fma(term(i), v[K][i], s);
    return s;
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
    out[t] = tensor_contract_fixed<N, D>(v, st, coef + l * cbs + base);
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

// ---------------- binned 3D evaluation ----------------
// With many scattered points per batch entry, a point's (D+1)^3 patch is
// (D+1)^2 separate rows of an array far beyond one XCD's L2.  The points are
// binned by tile of T^3 cells (binning.hpp); one workgroup per (batch entry,
// tile) stages the tile's (T+D)^3 coefficient patch and the knots and cell
// brackets of the non-uniform dimensions in LDS and evaluates its points
// from there, with the basis routines and the term order of
// evalnd_fixed_kernel: the values are bit-identical to it.
constexpr int kNdBinT = 256;      // threads of the count workgroups
constexpr int kNdBinPts = 16;     // points per thread there
constexpr int kNdTileThreads = 512;
constexpr size_t kNdPatchLds = 64 * 1024;
constexpr size_t kNdTabLds = 32 * 1024;

struct Bin3D {
    int T;
    int nt[3];
    int nbins;
};

// the LDS copies of the knots and cell brackets of the non-uniform
// dimensions: knots of every dimension, then brackets of every dimension
size_t nd_tables_bytes(NdArgs const& a, int N)
{
    size_t n = 0;
    for (int k = 0; k < N; ++k)
        if (!a.b[k].uniform)
            n += size_t(a.b[k].ncells + 1 + 2 * a.b[k].degree) * sizeof(double) + size_t(a.b[k].nbucket) * 4;
    return n;
}

template <int N>
__device__ void stage_tables_nd(NdArgs& a, double* lds, int tid, int nthreads)
{
    double* kn = lds;
    int nk_total = 0;
#pragma unroll
    for (int k = 0; k < N; ++k) nk_total += a.b[k].uniform ? 0 : a.b[k].ncells + 1 + 2 * a.b[k].degree;
    uint32_t* cl = reinterpret_cast<uint32_t*>(lds + nk_total);
#pragma unroll
    for (int k = 0; k < N; ++k) {
        BSplinesView& b = a.b[k];
        int const nk = b.uniform ? 0 : b.ncells + 1 + 2 * b.degree, nc = b.uniform ? 0 : b.nbucket;
        for (int i = tid; i < nk; i += nthreads) kn[i] = b.knots[i];
        for (int i = tid; i < nc; i += nthreads) cl[i] = b.cells[i];
        b.knots = kn;
        b.cells = cl;
        kn += nk;
        cl += nc;
    }
}

// SplineEvaluatorND::eval's point handling (evalnd_fixed_kernel): periodic
// wrap, then the Null rule (0, no tile) unless some order is nonzero
template <int N>
__device__ __forceinline__ bool nd_place(NdArgs const& a, double (&x)[N], int any_order)
{
    if (any_order) return true;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        BSplinesView const& b = a.b[k];
        if (b.periodic && (x[k] < b.rmin || x[k] > b.rmax)) x[k] -= floor((x[k] - b.rmin) / b.length) * b.length;
    }
#pragma unroll
    for (int k = 0; k < N; ++k) {
        BSplinesView const& b = a.b[k];
        if (!b.periodic && ((x[k] < b.rmin && a.lower[k] == DDCX_EXTRAP_NULL) ||
                            (x[k] > b.rmax && a.upper[k] == DDCX_EXTRAP_NULL)))
            return false;
    }
    return true;
}

// each point's tile and the per-tile counts; points the Null rule sends to 0
// get their value here
__global__ void __launch_bounds__(kNdBinT) ndbin_count_kernel(NdArgs a, Bin3D g, const double* __restrict__ coords,
                                                              double* __restrict__ out, int64_t npts, int64_t bpb,
                                                              int* __restrict__ counts, uint16_t* __restrict__ bins,
                                                              int any_order)
{
    __shared__ int hist[binning::kMaxBins];
    extern __shared__ double lds_tabs[];
    int64_t const l = blockIdx.x / bpb;
    int64_t const p0 = (blockIdx.x % bpb) * int64_t(kNdBinT * kNdBinPts);
    for (int k = threadIdx.x; k < g.nbins; k += kNdBinT) hist[k] = 0;
    stage_tables_nd<3>(a, lds_tabs, threadIdx.x, kNdBinT);
    __syncthreads();
    // the coordinates of kHold points are requested before the first is binned
    constexpr int kHold = 8;
    static_assert(kNdBinPts % kHold == 0, "points per thread in whole batches");
    for (int u0 = 0; u0 < kNdBinPts; u0 += kHold) {
        double xs[kHold][3];
#pragma unroll
        for (int u = 0; u < kHold; ++u) {
            int64_t const p = p0 + (u0 + u) * kNdBinT + threadIdx.x;
            int64_t const t = l * npts + (p < npts ? p : npts - 1);
#pragma unroll
            for (int k = 0; k < 3; ++k) xs[u][k] = coords[3 * t + k];
        }
#pragma unroll
        for (int u = 0; u < kHold; ++u) {
            int64_t const p = p0 + (u0 + u) * kNdBinT + threadIdx.x;
            if (p < npts) {
                int64_t const t = l * npts + p;
                double x[3] = {xs[u][0], xs[u][1], xs[u][2]};
                if (!nd_place<3>(a, x, any_order)) {
                    out[t] = 0.0;
                    bins[t] = uint16_t(binning::kNoBin);
                    continue;
                }
                int bin = 0;
#pragma unroll
                for (int k = 0; k < 3; ++k) bin = bin * g.nt[k] + cell_jmin(a.b[k], x[k]) / g.T;
                bins[t] = uint16_t(bin);
                atomicAdd(&hist[bin], 1);
            }
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < g.nbins; k += kNdBinT)
        if (hist[k]) atomicAdd(&counts[l * g.nbins + k], hist[k]);
}

template <int D, bool ANY>
__global__ void __launch_bounds__(kNdTileThreads, D <= 3 ? 4 : 1) ndbin_tile_kernel(NdArgs a, Bin3D g,
                                                                    const double* __restrict__ recs,
                                                                    double* __restrict__ out, int64_t npts,
                                                                    const double* __restrict__ coef, int64_t cbs,
                                                                    const int* __restrict__ starts,
                                                                    const int* __restrict__ perm)
{
    constexpr int N = 3;
    extern __shared__ double sh[];
    int const P = g.T + D;
    double* const cp = sh; // coefficient patch, P^3, dimension 3 fastest
    int64_t const l = blockIdx.x / g.nbins;
    int const bin = int(blockIdx.x % g.nbins);
    int a0[N];
    {
        int r = bin;
#pragma unroll
        for (int k = N - 1; k >= 0; --k) {
            a0[k] = (r % g.nt[k]) * g.T;
            r /= g.nt[k];
        }
    }
    int sz[N];
#pragma unroll
    for (int k = 0; k < N; ++k) sz[k] = a.b[k].ncells + a.b[k].degree;
    const double* c = coef + l * cbs;
    int const P3 = P * P * P;
    for (int idx = threadIdx.x; idx < P3; idx += kNdTileThreads) {
        int const i0 = idx / (P * P), i1 = (idx / P) % P, i2 = idx % P;
        bool const in = a0[0] + i0 < sz[0] && a0[1] + i1 < sz[1] && a0[2] + i2 < sz[2];
        cp[idx] = in ? c[int64_t(a0[0] + i0) * a.st[0] + int64_t(a0[1] + i1) * a.st[1] + int64_t(a0[2] + i2) * a.st[2]]
                     : 0.0;
    }
    stage_tables_nd<N>(a, sh + P3, threadIdx.x, kNdTileThreads);
    __syncthreads();
    int64_t const pst[N] = {int64_t(P) * P, P, 1};
    int const k0 = starts[l * (g.nbins + 1) + bin], k1 = starts[l * (g.nbins + 1) + bin + 1];
    // U records per thread and step, the next step's in flight while the
    // current ones are evaluated; 2 for derivatives, whose tables take more
    // registers: at most 128 VGPRs (the launch bound for D <= 3) keep two
    // workgroups per CU, as the LDS allows
    constexpr int U = ANY ? 2 : 4;
    double nx[U][N];
    int np[U];
    auto fetch = [&](int kb) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            int const k = min(kb + u * kNdTileThreads, k1 - 1); // past the end: the last record again, unused
            int64_t const q = l * npts + k;
#pragma unroll
            for (int d = 0; d < N; ++d) nx[u][d] = recs[3 * q + d];
            np[u] = perm[q];
        }
    };
    int const kbeg = k0 + int(threadIdx.x);
    if (kbeg < k1) fetch(kbeg);
    for (int kb = kbeg; kb < k1; kb += U * kNdTileThreads) {
        double cx[U][N];
        int cpn[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int d = 0; d < N; ++d) cx[u][d] = nx[u][d];
            cpn[u] = np[u];
        }
        if (kb + U * kNdTileThreads < k1) fetch(kb + U * kNdTileThreads);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (kb + u * kNdTileThreads >= k1) break;
            double x[N];
#pragma unroll
            for (int d = 0; d < N; ++d) x[d] = cx[u][d];
            nd_place<N>(a, x, ANY ? 1 : 0); // the wrap; every binned point passes the rules
            double v[N][kM];
            int64_t base = 0;
#pragma unroll
            for (int d = 0; d < N; ++d) {
                int const j = basis_fixed<D, ANY>(a.b[d], x[d], ANY ? a.order[d] : 0, v[d]);
                base += int64_t(j - a0[d]) * pst[d];
            }
            out[l * npts + cpn[u]] = tensor_contract_fixed<N, D>(v, pst, cp + base);
        }
    }
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

// tile side T (the patch of (T+D)^3 doubles fits kNdPatchLds) and tiles; none
// (nbins 0) when the grid is not 3D with one degree, the tiles exceed the
// histograms or the knots and brackets do not fit in LDS
Bin3D bins3d(NdArgs const& a, int nd)
{
    Bin3D g{};
    if (nd != 3) return g;
    int const D = a.b[0].degree;
    for (int k = 0; k < 3; ++k)
        if (a.b[k].degree != D || D > 5 || (!a.b[k].uniform && a.b[k].nbucket == 0)) return g;
    if (nd_tables_bytes(a, 3) > kNdTabLds) return g;
    int T = 16;
    while (T > 4 && size_t(T + D) * (T + D) * (T + D) * sizeof(double) > kNdPatchLds) --T;
    int64_t nbins = 1;
    for (int k = 0; k < 3; ++k) {
        g.nt[k] = (a.b[k].ncells + T - 1) / T;
        nbins *= g.nt[k];
    }
    if (nbins > binning::kMaxBins) return Bin3D{};
    g.T = T;
    g.nbins = int(nbins);
    return g;
}

} // namespace

int64_t evalnd_binned_workspace(BSplinesPlan const* const* bs, int nd, int64_t npts, int64_t nbatch)
{
    // sizes only: the views' shapes, no device access
    NdArgs a{};
    for (int k = 0; k < nd && k < kMaxNd; ++k) a.b[k] = bs[k]->device_view();
    Bin3D const g = bins3d(a, nd);
    if (g.nbins == 0 || npts >= (int64_t(1) << 31)) return 0;
    return binning::carve(g.nbins, 3, npts, nbatch, nullptr).bytes;
}

void launch_evalnd_binned(BSplinesPlan const* const* bs, int nd, const int* orders, const int* lower,
                          const int* upper, double* out, const double* coords, int64_t npts, const double* coef,
                          int64_t cbs, const int64_t* st, int64_t nbatch, void* work, int64_t work_bytes,
                          hipStream_t stream)
{
    using namespace binning;
    NdArgs const a = make_args(bs, nd, orders, lower, upper, st);
    Bin3D const g = bins3d(a, nd);
    int64_t const need = evalnd_binned_workspace(bs, nd, npts, nbatch);
    if (need == 0) throw Error(DDCX_ERR_UNSUPPORTED, "binned N-D evaluation: 3 dimensions of one degree <= 5 only");
    if (work_bytes < need) throw Error(DDCX_ERR_INVALID, "binned N-D evaluation: workspace too small");
    if (npts * nbatch == 0) return;
    int any = 0;
    for (int k = 0; k < nd; ++k) any |= a.order[k];
    Work const ws = carve(g.nbins, 3, npts, nbatch, work);
    size_t const tabs = nd_tables_bytes(a, 3);
    hipLaunchKernelGGL(fill_int_kernel, dim3(grid_blocks(nbatch * g.nbins, 256)), dim3(256), 0, stream, ws.counts,
                       nbatch * g.nbins);
    int64_t const bpb = (npts + kNdBinT * kNdBinPts - 1) / (kNdBinT * kNdBinPts);
    hipLaunchKernelGGL(ndbin_count_kernel, dim3(grid_blocks(bpb * nbatch, 1)), dim3(kNdBinT), tabs, stream, a, g,
                       coords, out, npts, bpb, ws.counts, ws.bins, any);
    scan_and_scatter<3, false>(g.nbins, coords, nullptr, npts, nbatch, ws, stream);
    bool launched = false;
    dispatch_degree(a.b[0].degree, [&]<int D>() {
        if constexpr (D <= 5) {
            size_t const lds = size_t(g.T + D) * (g.T + D) * (g.T + D) * sizeof(double) + tabs;
            auto go = [&](auto kern) {
                hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                              hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)),
                          "hipFuncSetAttribute");
                hipLaunchKernelGGL(kern, dim3(grid_blocks(nbatch * g.nbins, 1)), dim3(kNdTileThreads), lds, stream, a,
                                   g, ws.recs, out, npts, coef, cbs, ws.starts, ws.perm);
            };
            if (any) go(ndbin_tile_kernel<D, true>);
            else go(ndbin_tile_kernel<D, false>);
            launched = true;
        }
    });
    if (!launched) throw Error(DDCX_ERR_UNSUPPORTED, "binned N-D evaluation supports degrees 1..5");
    hip_check(hipGetLastError(), "binned N-D evaluation launch");
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
