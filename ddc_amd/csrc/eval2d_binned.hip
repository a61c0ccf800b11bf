// Adjoint of the batched 2D evaluation (value form) with the points binned by
// cell tile, instead of one global atomic per point and coefficient.
//
//   A  count:   each point's tile (T x T cells) -> per-tile counts (LDS
//               histogram per workgroup, one global add per tile it saw);
//               points that the Null rule sends to 0 get gx = 0 here
//   B  scan:    per batch entry, tile start offsets
//   C  scatter: each workgroup reserves one range per tile it saw and writes
//               its points' indices there (perm)
//   D  tiles:   one workgroup per (batch entry, tile) stages the tile's
//               coefficient patch ((T + d1) x (T + d2)) in LDS, evaluates
//               gx = w * grad f for its points from it, accumulates
//               w * B1 (x) B2 in an LDS gradient patch with LDS atomics and
//               adds the patch to gcoef (zeroed first) with one global atomic
//               per nonzero entry (neighbouring patches overlap by d rows)
//
// The point semantics (periodic wrap, Null and Constant rules, zero
// coordinate gradient along a clamped dimension) are those of
// eval2d_vjp_kernel.  The order of the additions inside a patch depends on
// scheduling, as with any atomic accumulation.
#include "internal.hpp"

namespace ddcx {

namespace {

constexpr int kBinT = 256;       // threads of the count / scatter workgroups
constexpr int kBinPts = 16;      // points per thread there (4096 per workgroup)
constexpr int kMaxBins = 4096;   // tiles per batch entry
constexpr int kTileThreads = 256;

struct Bin2D {
    int T;        // tile side in cells
    int nt1, nt2; // tiles per dimension
    int nbins;
};

// first B-spline index of the cell holding x (the jmin of eval_basis)
__device__ __forceinline__ int cell_of(BSplinesView const& b, double x)
{
    if (b.uniform) {
        double off;
        return uniform_cell(b, x, off);
    }
    return nonuniform_cell(b, b.knots, x) - b.degree;
}

// tile of point t, or -1 (Null rule: gx = 0 written here when `write_gx`)
__device__ __forceinline__ int bin_of(BSplinesView const& b1, BSplinesView const& b2, Rules2D rules, Bin2D const& g,
                                      const double* coords, double* gx, int64_t t, bool write_gx)
{
    double x1 = coords[2 * t], x2 = coords[2 * t + 1];
    bool in1, in2;
    if (!place2d(b1, b2, rules, x1, x2, in1, in2)) {
        if (write_gx) {
            gx[2 * t] = 0.0;
            gx[2 * t + 1] = 0.0;
        }
        return -1;
    }
    return (cell_of(b1, x1) / g.T) * g.nt2 + cell_of(b2, x2) / g.T;
}

__global__ void __launch_bounds__(256) fill_int_kernel(int* __restrict__ p, int64_t n)
{
    int64_t const i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
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
                                                          int64_t npts, int64_t bpb, int* __restrict__ counts)
{
    __shared__ int hist[kMaxBins];
    int64_t const l = blockIdx.x / bpb;
    int64_t const p0 = (blockIdx.x % bpb) * int64_t(kBinT * kBinPts);
    for (int k = threadIdx.x; k < g.nbins; k += kBinT) hist[k] = 0;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kBinPts; ++u) {
        int64_t const p = p0 + u * kBinT + threadIdx.x;
        if (p < npts) {
            int const bin = bin_of(b1, b2, rules, g, coords, gx, l * npts + p, true);
            if (bin >= 0) atomicAdd(&hist[bin], 1);
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < g.nbins; k += kBinT)
        if (hist[k]) atomicAdd(&counts[l * g.nbins + k], hist[k]);
}

// exclusive scan of one batch entry's tile counts (nbins <= 4096, 1024 threads,
// 4 tiles each); starts gets nbins + 1 entries, cursor a copy of the first nbins
__global__ void __launch_bounds__(1024) bin_scan_kernel(Bin2D g, const int* __restrict__ counts,
                                                        int* __restrict__ starts, int* __restrict__ cursor)
{
    __shared__ int part[1024];
    int64_t const l = blockIdx.x;
    int const t = threadIdx.x;
    int c[4];
    int sum = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        int const k = 4 * t + u;
        c[u] = k < g.nbins ? counts[l * g.nbins + k] : 0;
        sum += c[u];
    }
    part[t] = sum;
    __syncthreads();
    for (int s = 1; s < 1024; s *= 2) { // inclusive Hillis-Steele scan
        int const v = t >= s ? part[t - s] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - sum;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        int const k = 4 * t + u;
        if (k < g.nbins) {
            starts[l * (g.nbins + 1) + k] = run;
            cursor[l * g.nbins + k] = run;
        }
        run += c[u];
    }
    if (t == 1023) starts[l * (g.nbins + 1) + g.nbins] = part[1023];
}

__global__ void __launch_bounds__(kBinT) bin_scatter_kernel(BSplinesView b1, BSplinesView b2, Rules2D rules, Bin2D g,
                                                            const double* __restrict__ coords, int64_t npts,
                                                            int64_t bpb, int* __restrict__ cursor,
                                                            int* __restrict__ perm)
{
    __shared__ int hist[kMaxBins];
    __shared__ int base[kMaxBins];
    int64_t const l = blockIdx.x / bpb;
    int64_t const p0 = (blockIdx.x % bpb) * int64_t(kBinT * kBinPts);
    for (int k = threadIdx.x; k < g.nbins; k += kBinT) hist[k] = 0;
    __syncthreads();
    int bin[kBinPts], rank[kBinPts];
#pragma unroll
    for (int u = 0; u < kBinPts; ++u) {
        int64_t const p = p0 + u * kBinT + threadIdx.x;
        bin[u] = p < npts ? bin_of(b1, b2, rules, g, coords, nullptr, l * npts + p, false) : -1;
        rank[u] = bin[u] >= 0 ? atomicAdd(&hist[bin[u]], 1) : 0;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < g.nbins; k += kBinT)
        if (hist[k]) base[k] = atomicAdd(&cursor[l * g.nbins + k], hist[k]);
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kBinPts; ++u) {
        int64_t const p = p0 + u * kBinT + threadIdx.x;
        if (bin[u] >= 0) perm[l * npts + base[bin[u]] + rank[u]] = int(p);
    }
}

template <int D1, int D2>
__global__ void __launch_bounds__(kTileThreads) bin_tile_kernel(BSplinesView b1, BSplinesView b2, Rules2D rules, Bin2D g,
                                                                const double* __restrict__ coords,
                                                                const double* __restrict__ w, double* __restrict__ gx,
                                                                int64_t npts, const double* __restrict__ coef,
                                                                double* __restrict__ gcoef, int64_t cbs, int64_t st1,
                                                                int64_t st2, const int* __restrict__ starts,
                                                                const int* __restrict__ perm)
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
    __syncthreads();
    int const k0 = starts[l * (g.nbins + 1) + bin], k1 = starts[l * (g.nbins + 1) + bin + 1];
    for (int k = k0 + int(threadIdx.x); k < k1; k += kTileThreads) {
        int64_t const t = l * npts + perm[l * npts + k];
        double x1 = coords[2 * t], x2 = coords[2 * t + 1];
        double const wt = w[t];
        bool in1, in2;
        place2d(b1, b2, rules, x1, x2, in1, in2); // true for every binned point
        double e1[(D1 + 1) * 2], e2[(D2 + 1) * 2];
        int const j1 = eval_basis_and_n_derivs<D1, 1, true>(b1, b1.knots, x1, e1) - r0;
        int const j2 = eval_basis_and_n_derivs<D2, 1, true>(b2, b2.knots, x2, e2) - c0;
        double f1 = 0.0, f2 = 0.0;
#pragma unroll
        for (int i = 0; i < D1 + 1; ++i)
#pragma unroll
            for (int j = 0; j < D2 + 1; ++j) {
                int const o = (j1 + i) * P2 + (j2 + j);
                double const cij = cp[o];
                f1 += cij * e1[i * 2 + 1] * e2[j * 2];
                f2 += cij * e1[i * 2] * e2[j * 2 + 1];
                atomicAdd(&gp[o], wt * e1[i * 2] * e2[j * 2]);
            }
        gx[2 * t] = in1 ? wt * f1 : 0.0;
        gx[2 * t + 1] = in2 ? wt * f2 : 0.0;
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
// histograms hold).  Two (T + 5)^2 patches of doubles fit in LDS for T <= 64.
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

struct Work2D {
    int* counts;
    int* starts;
    int* cursor;
    int* perm;
    int64_t bytes;
};

Work2D carve(Bin2D const& g, int64_t npts, int64_t nbatch, void* base)
{
    auto up = [](int64_t n) { return (n * int64_t(sizeof(int)) + 255) / 256 * 256; };
    int64_t const a = up(nbatch * g.nbins), b = up(nbatch * (g.nbins + 1)), c = up(nbatch * g.nbins),
                  d = up(nbatch * npts);
    char* p = static_cast<char*>(base);
    return Work2D{reinterpret_cast<int*>(p), reinterpret_cast<int*>(p + a), reinterpret_cast<int*>(p + a + b),
                  reinterpret_cast<int*>(p + a + b + c), a + b + c + d};
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
    if (g.nbins == 0 || npts >= (int64_t(1) << 31) || bp1.view.degree > 5 || bp2.view.degree > 5) return 0;
    return carve(g, npts, nbatch, nullptr).bytes;
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
    Work2D const ws = carve(g, npts, nbatch, work);
    auto blocks = [](int64_t n, int t) {
        int64_t const b = (n + t - 1) / t;
        if (b > int64_t(0x7fffffff)) throw Error(DDCX_ERR_INVALID, "batch too large for one launch");
        return unsigned(b);
    };
    int const s1 = bp1.size(), s2 = bp2.size();
    hipLaunchKernelGGL(zero_patch_kernel, dim3(blocks(int64_t(s1) * s2 * nbatch, 256)), dim3(256), 0, stream, gcoef,
                       s1, s2, cbs, st1, st2, nbatch);
    hipLaunchKernelGGL(fill_int_kernel, dim3(blocks(nbatch * g.nbins, 256)), dim3(256), 0, stream, ws.counts,
                       nbatch * g.nbins);
    if (npts > 0) {
        int64_t const bpb = (npts + kBinT * kBinPts - 1) / (kBinT * kBinPts);
        hipLaunchKernelGGL(bin_count_kernel, dim3(blocks(bpb * nbatch, 1)), dim3(kBinT), 0, stream, b1, b2, rules, g,
                           coords, gx, npts, bpb, ws.counts);
        hipLaunchKernelGGL(bin_scan_kernel, dim3(unsigned(nbatch)), dim3(1024), 0, stream, g, ws.counts, ws.starts,
                           ws.cursor);
        hipLaunchKernelGGL(bin_scatter_kernel, dim3(blocks(bpb * nbatch, 1)), dim3(kBinT), 0, stream, b1, b2, rules, g,
                           coords, npts, bpb, ws.cursor, ws.perm);
        bool launched = false;
        dispatch_degree(b1.degree, [&]<int D1>() {
            dispatch_degree(b2.degree, [&]<int D2>() {
                if constexpr (D1 <= 5 && D2 <= 5) {
                    size_t const lds = size_t(2) * (g.T + D1) * (g.T + D2) * sizeof(double);
                    auto kern = bin_tile_kernel<D1, D2>;
                    hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                                  hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)),
                              "hipFuncSetAttribute");
                    hipLaunchKernelGGL(kern, dim3(blocks(nbatch * g.nbins, 1)), dim3(kTileThreads), lds, stream, b1,
                                       b2, rules, g, coords, w, gx, npts, coef, gcoef, cbs, st1, st2, ws.starts, ws.perm);
                    launched = true;
                }
            });
        });
        if (!launched) throw Error(DDCX_ERR_UNSUPPORTED, "binned 2D adjoint supports degrees 1..5");
    }
    hip_check(hipGetLastError(), "binned 2D adjoint launch");
}

} // namespace ddcx
