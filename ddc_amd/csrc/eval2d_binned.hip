// Adjoint of the batched 2D evaluation (value form) with the points binned by
// cell tile, instead of one global atomic per point and coefficient.
//
//   A  count:   each point's tile (T x T cells) -> per-tile counts (LDS
//               histogram per workgroup, one global add per tile it saw),
//               and the tile itself (16 bits per point, for C); points that
//               the Null rule sends to 0 get gx = 0 here
//   B  scan:    per batch entry, tile start offsets
//   C  scatter: each workgroup reserves one range per tile it saw and writes
//               its points there in tile order: index (perm) and the record
//               (x1, x2, w)
//   D  tiles:   one workgroup per (batch entry, tile) stages the tile's
//               coefficient patch ((T + d1) x (T + d2)) in LDS, reads its
//               points' records in order, evaluates gx = w * grad f from
//               the patch (stored through perm), accumulates
//               w * B1 (x) B2 in an LDS gradient patch with LDS atomics and
//               adds the patch to gcoef (zeroed first) with one global atomic
//               per nonzero entry (neighbouring patches overlap by d rows)
//
// The knots and cell brackets of non-uniform dimensions are staged in LDS by
// every kernel that locates points (grids whose tables do not fit take the
// one-atomic-per-term kernel instead: the workspace query returns 0).
// The point semantics (periodic wrap, Null and Constant rules, zero
// coordinate gradient along a clamped dimension) are those of
// eval2d_vjp_kernel.  The order of the additions inside a patch depends on
// scheduling, as with any atomic accumulation.
#include "binning.hpp"

namespace ddcx {

namespace {

using namespace binning;

constexpr int kBinT = 256;       // threads of the count workgroups
constexpr int kBinPts = 16;      // points per thread there (4096 per workgroup)
constexpr int kTileThreads = 1024;
constexpr int kTileRecs = 4;    // records per thread in flight in the tile kernel
constexpr size_t kTabLds = 48 * 1024; // LDS for staged knots and brackets

struct Bin2D {
    int T;        // tile side in cells
    int nt1, nt2; // tiles per dimension
    int nbins;
};

// tile of the point (x1, x2) as given (before wrap and rules), or -1 when the
// Null rule sends it to 0
__device__ __forceinline__ int bin_of_xy(BSplinesView const& b1, BSplinesView const& b2, Rules2D rules,
                                         Bin2D const& g, double x1, double x2)
{
    bool in1, in2;
    if (!place2d(b1, b2, rules, x1, x2, in1, in2)) return -1;
    return (cell_jmin(b1, x1) / g.T) * g.nt2 + cell_jmin(b2, x2) / g.T;
}

// tile of point t, or -1 (Null rule: gx = 0 written here when `write_gx`)
__device__ __forceinline__ int bin_of(BSplinesView const& b1, BSplinesView const& b2, Rules2D rules, Bin2D const& g,
                                      const double* coords, double* gx, int64_t t, bool write_gx)
{
    int const bin = bin_of_xy(b1, b2, rules, g, coords[2 * t], coords[2 * t + 1]);
    if (bin < 0 && write_gx) {
        gx[2 * t] = 0.0;
        gx[2 * t + 1] = 0.0;
    }
    return bin;
}

__global__ void __launch_bounds__(256) zero_patch_kernel(double* __restrict__ g, int s1, int s2, int64_t cbs,
                                                         int64_t st1, int64_t st2, int64_t nbatch)
{
    int64_t const t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    int64_t const per = int64_t(s1) * s2;
    if (t >= per * nbatch) return;
    int64_t const l = t / per, r = t % per;
    g[l * cbs + (r / s2) * st1 + (r % s2) * st2] = 0.0;
}

__global__ void __launch_bounds__(kBinT) bin_count_kernel(BSplinesView b1, BSplinesView b2, Rules2D rules, Bin2D g,
                                                          const double* __restrict__ coords, double* __restrict__ gx,
                                                          int64_t npts, int64_t bpb, int* __restrict__ counts,
                                                          uint16_t* __restrict__ bins)
{
    __shared__ int hist[kMaxBins];
    extern __shared__ double lds_tabs[];
    int64_t const l = blockIdx.x / bpb;
    int64_t const p0 = (blockIdx.x % bpb) * int64_t(kBinT * kBinPts);
    for (int k = threadIdx.x; k < g.nbins; k += kBinT) hist[k] = 0;
    stage_tables_2d(b1, b2, lds_tabs, threadIdx.x, kBinT);
    __syncthreads();
    // every point's coordinates are requested before the first is binned
    double2 xy[kBinPts];
#pragma unroll
    for (int u = 0; u < kBinPts; ++u) {
        int64_t const p = p0 + u * kBinT + threadIdx.x;
        xy[u] = reinterpret_cast<const double2*>(coords)[l * npts + (p < npts ? p : npts - 1)];
    }
#pragma unroll
    for (int u = 0; u < kBinPts; ++u) {
        int64_t const p = p0 + u * kBinT + threadIdx.x;
        if (p < npts) {
            int64_t const t = l * npts + p;
            int const bin = bin_of_xy(b1, b2, rules, g, xy[u].x, xy[u].y);
            if (bin < 0) {
                gx[2 * t] = 0.0;
                gx[2 * t + 1] = 0.0;
            }
            bins[t] = uint16_t(bin < 0 ? kNoBin : bin);
            if (bin >= 0) atomicAdd(&hist[bin], 1);
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < g.nbins; k += kBinT)
        if (hist[k]) atomicAdd(&counts[l * g.nbins + k], hist[k]);
}

template <int D1, int D2>
__global__ void __launch_bounds__(kTileThreads) bin_tile_kernel(BSplinesView b1, BSplinesView b2, Rules2D rules, Bin2D g,
                                                                const double* __restrict__ recs,
                                                                double* __restrict__ gx, int64_t npts,
                                                                const double* __restrict__ coef,
                                                                double* __restrict__ gcoef, int64_t cbs, int64_t st1,
                                                                int64_t st2, const int* __restrict__ starts,
                                                                const int* __restrict__ perm, int stage)
{
    extern __shared__ double sh[];
    int const P1 = g.T + D1, P2 = g.T + D2;
    double* cp = sh;           // coefficient patch
    double* gp = sh + P1 * P2; // gradient patch
    int64_t const l = blockIdx.x / g.nbins;
    int const bin = int(blockIdx.x % g.nbins);
    int const r0 = (bin / g.nt2) * g.T, c0 = (bin % g.nt2) * g.T;
    int const s1 = b1.ncells + b1.degree, s2 = b2.ncells + b2.degree;
    const double* c = coef + l * cbs;
    for (int k = threadIdx.x; k < P1 * P2; k += kTileThreads) {
        int const a = k / P2, b = k % P2;
        bool const in = r0 + a < s1 && c0 + b < s2;
        cp[k] = in ? c[int64_t(r0 + a) * st1 + int64_t(c0 + b) * st2] : 0.0;
        gp[k] = 0.0;
    }
    // The knots and brackets are staged in LDS behind a run-time flag (set by
    // the launcher), so their reads compile to generic (flat) loads: measured
    // faster in this kernel than LDS instructions, which share the LDS pipe
    // with the 32 patch accesses per point, and than global loads.
    if (stage) stage_tables_2d(b1, b2, sh + 2 * P1 * P2, threadIdx.x, kTileThreads);
    __syncthreads();
    int const k0 = starts[l * (g.nbins + 1) + bin], k1 = starts[l * (g.nbins + 1) + bin + 1];
    // kTileRecs records per thread and step, the next step's in flight while
    // the current ones are processed (one record per thread left the kernel
    // waiting on memory latency)
    constexpr int U = D1 + D2 <= 7 ? kTileRecs : 2; // high degrees: fewer, to stay within the registers
    double n1[U], n2[U], nw[U];
    int np[U];
    auto fetch = [&](int kb) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            int const k = min(kb + u * kTileThreads, k1 - 1); // past the end: the last record again, unused
            int64_t const q = l * npts + k;
            n1[u] = recs[3 * q];
            n2[u] = recs[3 * q + 1];
            nw[u] = recs[3 * q + 2];
            np[u] = perm[q];
        }
    };
    int const kbeg = k0 + int(threadIdx.x);
    if (kbeg < k1) fetch(kbeg);
    for (int kb = kbeg; kb < k1; kb += U * kTileThreads) {
        double c1[U], c2[U], cw[U];
        int cpn[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            c1[u] = n1[u];
            c2[u] = n2[u];
            cw[u] = nw[u];
            cpn[u] = np[u];
        }
        if (kb + U * kTileThreads < k1) fetch(kb + U * kTileThreads);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (kb + u * kTileThreads >= k1) break;
            double x1 = c1[u], x2 = c2[u];
            double const wt = cw[u];
            int64_t const t = l * npts + cpn[u];
            bool in1, in2;
            place2d(b1, b2, rules, x1, x2, in1, in2); // true for every binned point
            double e1[(D1 + 1) * 2], e2[(D2 + 1) * 2];
            int const j1 = eval_basis_and_n_derivs<D1, 1, true>(b1, b1.knots, x1, e1) - r0;
            int const j2 = eval_basis_and_n_derivs<D2, 1, true>(b2, b2.knots, x2, e2) - c0;
            // f1 = sum_i B1'_i sum_j c_ij B2_j, f2 = sum_i B1_i sum_j c_ij B2'_j
            double f1 = 0.0, f2 = 0.0;
#pragma unroll
            for (int i = 0; i < D1 + 1; ++i) {
                double s0 = 0.0, s1 = 0.0;
                double const we = wt * e1[i * 2];
#pragma unroll
                for (int j = 0; j < D2 + 1; ++j) {
                    int const o = (j1 + i) * P2 + (j2 + j);
                    double const cij = cp[o];
                    s0 = fma(cij, e2[j * 2], s0);
                    s1 = fma(cij, e2[j * 2 + 1], s1);
                    atomicAdd(&gp[o], we * e2[j * 2]);
                }
                f1 = fma(e1[i * 2 + 1], s0, f1);
                f2 = fma(e1[i * 2], s1, f2);
            }
            gx[2 * t] = in1 ? wt * f1 : 0.0;
            gx[2 * t + 1] = in2 ? wt * f2 : 0.0;
        }
    }
    __syncthreads();
    double* gc = gcoef + l * cbs;
    for (int k = threadIdx.x; k < P1 * P2; k += kTileThreads) {
        double const v = gp[k];
        int const a = k / P2, b = k % P2;
        if (v != 0.0 && r0 + a < s1 && c0 + b < s2) atomicAdd(&gc[int64_t(r0 + a) * st1 + int64_t(c0 + b) * st2], v);
    }
}

// Tiles of 64 x 64 cells (few tiles per batch entry, so the count and scatter
// workgroups make few global additions), or 32 x 32 when 64 would leave fewer
// than 2048 tile workgroups in the launch (or more tiles than the LDS
// histograms hold).  Two (T + 5)^2 patches of doubles and the staged tables
// fit in LDS for T <= 64.
Bin2D bins_for(BSplinesPlan const& bp1, BSplinesPlan const& bp2, int64_t nbatch)
{
    auto make = [&](int T) {
        Bin2D g{};
        g.T = T;
        g.nt1 = (bp1.view.ncells + T - 1) / T;
        g.nt2 = (bp2.view.ncells + T - 1) / T;
        g.nbins = g.nt1 * g.nt2;
        return g;
    };
    Bin2D const g64 = make(64), g32 = make(32);
    if (g64.nbins <= kMaxBins && (int64_t(g64.nbins) * nbatch >= 2048 || g32.nbins > kMaxBins)) return g64;
    if (g32.nbins <= kMaxBins) return g32;
    Bin2D none{};
    return none;
}

} // namespace

// The workspace of the allocation-free entry point
// (ddcx_evaluator2d_eval_vjp_binned_static): a file-scope device array, one
// copy per GPU, held until the process ends.  Host code never uses the
// array's name as an address: static_scratch() resolves each device's copy
// once with hipGetSymbolAddress and keeps it.
static __device__ __align__(256) unsigned char scratch[DDCX_STATIC_SCRATCH_BYTES];

void* static_scratch()
{
    static std::mutex m;
    static std::vector<void*> ptrs;
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lock(m);
    if (int(ptrs.size()) <= dev) ptrs.resize(size_t(dev) + 1, nullptr);
    if (!ptrs[size_t(dev)]) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        ptrs[size_t(dev)] = p;
    }
    return ptrs[size_t(dev)];
}

int64_t eval2d_vjp_workspace(BSplinesPlan const& bp1, BSplinesPlan const& bp2, int64_t npts, int64_t nbatch)
{
    Bin2D const g = bins_for(bp1, bp2, nbatch);
    if (g.nbins == 0 || npts >= (int64_t(1) << 31) || bp1.view.degree > 5 || bp2.view.degree > 5 ||
        !staged_tables_fit_2d(bp1.device_view(), bp2.device_view(), kTabLds))
        return 0;
    return carve(g.nbins, 3, npts, nbatch, nullptr).bytes;
}

void launch_eval2d_vjp_binned(BSplinesPlan const& bp1, BSplinesPlan const& bp2, Rules2D const& rules,
                              const double* coords, const double* w, double* gx, int64_t npts, const double* coef,
                              double* gcoef, int64_t cbs, int64_t st1, int64_t st2, int64_t nbatch, void* work,
                              int64_t work_bytes, hipStream_t stream)
{
    bp1.ensure_device();
    bp2.ensure_device();
    BSplinesView const b1 = bp1.device_view(), b2 = bp2.device_view();
    Bin2D const g = bins_for(bp1, bp2, nbatch);
    int64_t const need = eval2d_vjp_workspace(bp1, bp2, npts, nbatch);
    if (need == 0) throw Error(DDCX_ERR_UNSUPPORTED, "binned 2D adjoint: grid or degree out of range");
    if (work_bytes < need) throw Error(DDCX_ERR_INVALID, "binned 2D adjoint: workspace too small");
    Work const ws = carve(g.nbins, 3, npts, nbatch, work);
    auto blocks = grid_blocks;
    int const s1 = bp1.size(), s2 = bp2.size();
    hipLaunchKernelGGL(zero_patch_kernel, dim3(blocks(int64_t(s1) * s2 * nbatch, 256)), dim3(256), 0, stream, gcoef,
                       s1, s2, cbs, st1, st2, nbatch);
    hipLaunchKernelGGL(fill_int_kernel, dim3(blocks(nbatch * g.nbins, 256)), dim3(256), 0, stream, ws.counts,
                       nbatch * g.nbins);
    if (npts > 0) {
        int64_t const bpb = (npts + kBinT * kBinPts - 1) / (kBinT * kBinPts);
        size_t const tabs = staged_tables_bytes_2d(b1, b2);
        hipLaunchKernelGGL(bin_count_kernel, dim3(blocks(bpb * nbatch, 1)), dim3(kBinT), tabs, stream, b1, b2, rules,
                           g, coords, gx, npts, bpb, ws.counts, ws.bins);
        scan_and_scatter<2, true>(g.nbins, coords, w, npts, nbatch, ws, stream);
        bool launched = false;
        dispatch_degree(b1.degree, [&]<int D1>() {
            dispatch_degree(b2.degree, [&]<int D2>() {
                if constexpr (D1 <= 5 && D2 <= 5) {
                    size_t const lds = size_t(2) * (g.T + D1) * (g.T + D2) * sizeof(double) + tabs;
                    auto kern = bin_tile_kernel<D1, D2>;
                    hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                                  hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)),
                              "hipFuncSetAttribute");
                    hipLaunchKernelGGL(kern, dim3(blocks(nbatch * g.nbins, 1)), dim3(kTileThreads), lds, stream, b1,
                                       b2, rules, g, ws.recs, gx, npts, coef, gcoef, cbs, st1, st2, ws.starts, ws.perm,
                                       1);
                    launched = true;
                }
            });
        });
        if (!launched) throw Error(DDCX_ERR_UNSUPPORTED, "binned 2D adjoint supports degrees 1..5");
    }
    hip_check(hipGetLastError(), "binned 2D adjoint launch");
}

} // namespace ddcx
