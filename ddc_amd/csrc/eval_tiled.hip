// LDS-staged batched evaluation (SplineEvaluator::operator() / deriv(),
// spline_evaluator.hpp:263-478).
//
// Random evaluation points gather the coefficients of their line at random,
// so a workgroup takes LB consecutive lines, stages their coefficient tile
// (size() x LB) and, for non-uniform bases, the knot vector and the cell-search
// brackets in LDS, and then evaluates all the points of those lines with
// coalesced coordinate loads and output stores.  HBM sees every coefficient,
// coordinate and output once.  Two thread mappings:
//  - batch-contiguous (LINE_MAJOR = false): lanes take consecutive lines, so
//    tile rows and point rows are read as LB*8-byte runs;
//  - interpolation-contiguous (LINE_MAJOR = true): the tile is read line by
//    line (size() contiguous doubles each) and the NT/LB threads of a line
//    take consecutive points of it.
#include "internal.hpp"

namespace ddcx {

namespace {

__device__ __forceinline__ int64_t lay(Layout const& L, int64_t r, int64_t l)
{
    return r * L.rs + (l % L.inner) * L.is + (l / L.inner) * L.os;
}

template <int D, int ORDER>
__device__ __forceinline__ double point_value(BSplinesView const& b, const double* kn, int lower, int upper, double x,
                                              const double* tile, int pitch)
{
    double v[D + 1];
    int jmin;
    if constexpr (ORDER == 0) {
        if (b.periodic) {
            if (x < b.rmin || x > b.rmax) x -= floor((x - b.rmin) / b.length) * b.length;
        } else if (x < b.rmin) {
            if (lower == DDCX_EXTRAP_NULL) return 0.0;
            x = b.rmin;
        } else if (x > b.rmax) {
            if (upper == DDCX_EXTRAP_NULL) return 0.0;
            x = b.rmax;
        }
        jmin = eval_basis<D, true>(b, kn, x, v);
    } else {
        double der[(D + 1) * (ORDER + 1)];
        jmin = eval_basis_and_n_derivs<D, ORDER, true>(b, kn, x, der);
#pragma unroll
        for (int i = 0; i < D + 1; ++i) v[i] = der[i * (ORDER + 1) + ORDER];
    }
    double y = 0.0;
#pragma unroll
    for (int i = 0; i < D + 1; ++i) y += tile[(jmin + i) * pitch] * v[i];
    return y;
}

template <int D, int ORDER, int LB, int NT, bool LINE_MAJOR>
__global__ void __launch_bounds__(NT) eval_tiled_kernel(BSplinesView b, int lower, int upper, double* __restrict__ out,
                                                        const double* __restrict__ coords, Layout pl, int64_t npts,
                                                        const double* __restrict__ coef, Layout cl, int64_t nbatch,
                                                        int ncoef, int pitch, int nknots)
{
    extern __shared__ double smem[];
    double* tile = smem;
    double* kn = smem + size_t(ncoef) * pitch;
    uint32_t* cells = reinterpret_cast<uint32_t*>(kn + nknots);
    constexpr int ROWS = NT / LB;
    constexpr int U = 8;           // points per pipeline stage per thread
    constexpr int TC = 16;         // tile rows per thread per load chunk
    int const tid = threadIdx.x;
    int const ln = LINE_MAJOR ? tid / ROWS : tid % LB;
    int const prow = LINE_MAJOR ? tid % ROWS : tid / LB;
    int64_t const line0 = int64_t(blockIdx.x) * LB;
    int64_t const line = line0 + ln;
    bool const active = line < nbatch;
    int64_t const lc = active ? line : 0;
    int64_t const cb = lay(cl, 0, lc);
    int64_t const pb = lay(pl, 0, lc);
    // first batch of coordinates, then the whole coefficient tile: every load
    // of the tile is in flight before the first LDS write
    double xa[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        int64_t const ii = prow + int64_t(u) * ROWS;
        xa[u] = ii < npts ? coords[pb + ii * pl.rs] : 0.0;
    }
    // the tile is loaded in chunks of TC elements per thread, every load of a
    // chunk in flight before its LDS writes
    if constexpr (LINE_MAJOR) {
        int const nl = int(min<int64_t>(LB, nbatch - line0));
        int const nel = nl * ncoef;
        for (int e0 = tid; e0 < nel; e0 += TC * NT) {
            double tv[TC];
#pragma unroll
            for (int k = 0; k < TC; ++k) {
                int const e = e0 + k * NT;
                int const l = e / ncoef, r = e - l * ncoef;
                tv[k] = e < nel ? coef[lay(cl, r, line0 + l)] : 0.0;
            }
#pragma unroll
            for (int k = 0; k < TC; ++k) {
                int const e = e0 + k * NT;
                int const l = e / ncoef, r = e - l * ncoef;
                if (e < nel) tile[r * pitch + l] = tv[k];
            }
        }
    } else {
        for (int c0 = prow; c0 < ncoef; c0 += TC * ROWS) {
            double tv[TC];
#pragma unroll
            for (int k = 0; k < TC; ++k) {
                int const r = c0 + k * ROWS;
                tv[k] = r < ncoef ? coef[cb + int64_t(r) * cl.rs] : 0.0;
            }
#pragma unroll
            for (int k = 0; k < TC; ++k) {
                int const r = c0 + k * ROWS;
                if (r < ncoef) tile[r * pitch + ln] = tv[k];
            }
        }
    }
    if (!b.uniform) {
        for (int k = tid; k < nknots; k += NT) kn[k] = b.knots[k];
        for (int k = tid; k < b.nbucket; k += NT) cells[k] = b.cells[k];
    }
    __syncthreads();
    if (!active) return;
    const double* knots = b.uniform ? nullptr : kn;
    BSplinesView bl = b;
    bl.cells = b.nbucket > 0 ? cells : nullptr;
    const double* tl = tile + ln;
    // software pipeline: the next batch of coordinates is in flight while the
    // current one is evaluated and stored
    for (int64_t i = prow; i < npts; i += U * ROWS) {
        double xb[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            int64_t const ii = i + int64_t(U + u) * ROWS;
            xb[u] = ii < npts ? coords[pb + ii * pl.rs] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            int64_t const ii = i + int64_t(u) * ROWS;
            if (ii < npts) out[pb + ii * pl.rs] = point_value<D, ORDER>(bl, knots, lower, upper, xa[u], tl, pitch);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) xa[u] = xb[u];
    }
}

// Adjoint of the value form on the same tiles: gx = w * f'(x) where the
// spline itself is evaluated (0 where an extrapolation rule applies), and the
// coefficient gradient of the LB lines accumulated in an LDS tile with LDS
// atomics, then written once (every row, so no separate zeroing).
template <int D, int LB, int NT, bool LINE_MAJOR>
__global__ void __launch_bounds__(NT) eval_vjp_tiled_kernel(BSplinesView b, int lower, int upper,
                                                            const double* __restrict__ coords,
                                                            const double* __restrict__ w, double* __restrict__ gx,
                                                            Layout pl, int64_t npts, const double* __restrict__ coef,
                                                            double* __restrict__ gcoef, Layout cl, int64_t nbatch,
                                                            int ncoef, int pitch, int nknots)
{
    extern __shared__ double smem[];
    double* tile = smem;
    double* gt = smem + size_t(ncoef) * pitch;
    double* kn = gt + size_t(ncoef) * pitch;
    uint32_t* cells = reinterpret_cast<uint32_t*>(kn + nknots);
    constexpr int ROWS = NT / LB;
    constexpr int U = 4;
    constexpr int TC = 16;
    int const tid = threadIdx.x;
    int const ln = LINE_MAJOR ? tid / ROWS : tid % LB;
    int const prow = LINE_MAJOR ? tid % ROWS : tid / LB;
    int64_t const line0 = int64_t(blockIdx.x) * LB;
    int64_t const line = line0 + ln;
    bool const active = line < nbatch;
    int64_t const lc = active ? line : 0;
    int64_t const cb = lay(cl, 0, lc);
    int64_t const pb = lay(pl, 0, lc);
    int const nl = int(min<int64_t>(LB, nbatch - line0));
    int const nel = nl * ncoef;
    // coefficient tile (as in eval_tiled_kernel) and a zeroed gradient tile
    if constexpr (LINE_MAJOR) {
        for (int e0 = tid; e0 < nel; e0 += TC * NT) {
            double tv[TC];
#pragma unroll
            for (int k = 0; k < TC; ++k) {
                int const e = e0 + k * NT;
                int const l = e / ncoef, r = e - l * ncoef;
                tv[k] = e < nel ? coef[lay(cl, r, line0 + l)] : 0.0;
            }
#pragma unroll
            for (int k = 0; k < TC; ++k) {
                int const e = e0 + k * NT;
                int const l = e / ncoef, r = e - l * ncoef;
                if (e < nel) tile[r * pitch + l] = tv[k];
            }
        }
    } else {
        for (int c0 = prow; c0 < ncoef; c0 += TC * ROWS) {
            double tv[TC];
#pragma unroll
            for (int k = 0; k < TC; ++k) {
                int const r = c0 + k * ROWS;
                tv[k] = r < ncoef ? coef[cb + int64_t(r) * cl.rs] : 0.0;
            }
#pragma unroll
            for (int k = 0; k < TC; ++k) {
                int const r = c0 + k * ROWS;
                if (r < ncoef) tile[r * pitch + ln] = tv[k];
            }
        }
    }
    for (int k = tid; k < ncoef * pitch; k += NT) gt[k] = 0.0;
    if (!b.uniform) {
        for (int k = tid; k < nknots; k += NT) kn[k] = b.knots[k];
        for (int k = tid; k < b.nbucket; k += NT) cells[k] = b.cells[k];
    }
    __syncthreads();
    const double* knots = b.uniform ? nullptr : kn;
    BSplinesView bl = b;
    bl.cells = b.nbucket > 0 ? cells : nullptr;
    if (active) {
        const double* tl = tile + ln;
        double* gl = gt + ln;
        for (int64_t i = prow; i < npts; i += U * ROWS) {
            double xa[U], wa[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                int64_t const ii = i + int64_t(u) * ROWS;
                xa[u] = ii < npts ? coords[pb + ii * pl.rs] : 0.0;
                wa[u] = ii < npts ? w[pb + ii * pl.rs] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                int64_t const ii = i + int64_t(u) * ROWS;
                if (ii >= npts) continue;
                double x = xa[u];
                bool outside;
                if (!apply_bc(bl, lower, upper, x, outside)) {
                    gx[pb + ii * pl.rs] = 0.0;
                    continue;
                }
                double der[(D + 1) * 2];
                int const jmin = eval_basis_and_n_derivs<D, 1, true>(bl, knots, x, der);
                double dy = 0.0;
#pragma unroll
                for (int k = 0; k < D + 1; ++k) {
                    dy += tl[(jmin + k) * pitch] * der[k * 2 + 1];
                    atomicAdd(&gl[(jmin + k) * pitch], wa[u] * der[k * 2]);
                }
                gx[pb + ii * pl.rs] = outside ? 0.0 : wa[u] * dy;
            }
        }
    }
    __syncthreads();
    if constexpr (LINE_MAJOR) {
        for (int e = tid; e < nel; e += NT) {
            int const l = e / ncoef, r = e - l * ncoef;
            gcoef[lay(cl, r, line0 + l)] = gt[r * pitch + l];
        }
    } else if (active) {
        for (int r = prow; r < ncoef; r += ROWS) gcoef[cb + int64_t(r) * cl.rs] = gt[r * pitch + ln];
    }
}

constexpr size_t kMaxLds = 160 * 1024;

template <int D, int ORDER>
bool launch_orders(BSplinesPlan const& bs, int lower, int upper, double* out, const double* coords, Layout pl,
                   int64_t npts, const double* coef, Layout cl, int64_t nbatch, hipStream_t stream)
{
    int const ncoef = bs.size();
    int const nknots = bs.view.uniform ? 0 : int(bs.knots_host.size());
    BSplinesView const b = bs.device_view();
    size_t const extra = nknots * sizeof(double) + size_t(b.nbucket) * sizeof(uint32_t);
    auto fits = [&](int lb, int pitch) { return size_t(ncoef) * pitch * sizeof(double) + extra <= kMaxLds; };
    bool const line_major = pl.rs == 1;
    auto go = [&]<int LB>(int pitch) {
        constexpr int NT = 1024;
        size_t const lds = size_t(ncoef) * pitch * sizeof(double) + extra;
        auto kern = line_major ? eval_tiled_kernel<D, ORDER, LB, NT, true> : eval_tiled_kernel<D, ORDER, LB, NT, false>;
        hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)),
                  "hipFuncSetAttribute");
        int64_t const blocks = (nbatch + LB - 1) / LB;
        hipLaunchKernelGGL(kern, dim3(unsigned(blocks)), dim3(NT), lds, stream, b, lower, upper, out, coords, pl, npts,
                           coef, cl, nbatch, ncoef, pitch, nknots);
        hip_check(hipGetLastError(), "eval_tiled_kernel launch");
        return true;
    };
    if (ncoef > 2560) return false;
    if (fits(16, 17)) return go.template operator()<16>(17);
    if (fits(16, 16)) return go.template operator()<16>(16);
    if (fits(8, 9)) return go.template operator()<8>(9);
    if (fits(8, 8)) return go.template operator()<8>(8);
    return false;
}

} // namespace

bool launch_eval_tiled(BSplinesPlan const& bs, int lower, int upper, int order, double* out, const double* coords,
                       Layout pl, int64_t npts, const double* coef, Layout cl, int64_t nbatch, hipStream_t stream)
{
    // tiny batches, and mixed layouts (points and coefficients not both
    // batch-contiguous or both interpolation-contiguous), take the generic kernel
    if (npts < 64 || nbatch < 8) return false;
    if ((pl.rs == 1) != (cl.rs == 1)) return false;
    bool launched = false;
    dispatch_degree(bs.view.degree, [&]<int D>() {
        auto ord = [&]<int O>() {
            if constexpr (O <= D) {
                if (!launched && order == O)
                    launched = launch_orders<D, O>(bs, lower, upper, out, coords, pl, npts, coef, cl, nbatch, stream);
            }
        };
        ord.template operator()<0>();
        ord.template operator()<1>();
        ord.template operator()<2>();
        ord.template operator()<3>();
    });
    return launched;
}

bool launch_eval_vjp_tiled(BSplinesPlan const& bs, int lower, int upper, const double* coords, const double* w,
                           double* gx, Layout pl, int64_t npts, const double* coef, double* gcoef, Layout cl,
                           int64_t nbatch, hipStream_t stream)
{
    if (npts < 64 || nbatch < 8) return false;
    if ((pl.rs == 1) != (cl.rs == 1)) return false;
    int const ncoef = bs.size();
    if (ncoef > 2560) return false;
    int const nknots = bs.view.uniform ? 0 : int(bs.knots_host.size());
    BSplinesView const b = bs.device_view();
    size_t const extra = nknots * sizeof(double) + size_t(b.nbucket) * sizeof(uint32_t);
    auto lds_of = [&](int pitch) { return 2 * size_t(ncoef) * pitch * sizeof(double) + extra; };
    int pitch = 0;
    if (lds_of(9) <= kMaxLds) pitch = 9;
    else if (lds_of(8) <= kMaxLds) pitch = 8;
    else return false;
    bool const line_major = pl.rs == 1;
    bool launched = false;
    dispatch_degree(b.degree, [&]<int D>() {
        constexpr int LB = 8, NT = 1024;
        auto kern = line_major ? eval_vjp_tiled_kernel<D, LB, NT, true> : eval_vjp_tiled_kernel<D, LB, NT, false>;
        size_t const lds = lds_of(pitch);
        hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)),
                  "hipFuncSetAttribute");
        int64_t const blocks = (nbatch + LB - 1) / LB;
        hipLaunchKernelGGL(kern, dim3(unsigned(blocks)), dim3(NT), lds, stream, b, lower, upper, coords, w, gx, pl,
                           npts, coef, gcoef, cl, nbatch, ncoef, pitch, nknots);
        hip_check(hipGetLastError(), "eval_vjp_tiled_kernel launch");
        launched = true;
    });
    return launched;
}

} // namespace ddcx
