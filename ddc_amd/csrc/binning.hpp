// Binning of scattered evaluation points by cell tile, shared by the binned
// 2D adjoint (eval2d_binned.hip) and the binned 3D evaluation
// (eval_nd_binned.hip).  A caller's count pass gives each point a tile (16
// bits, kNoBin for a point an extrapolation rule settles by itself) and per
// batch entry and tile a count; then
//   bin_scan_kernel    turns the counts into tile offsets (starts) and a copy
//                      of them (cursor) for the scatter;
//   bin_scatter_kernel writes each point's record (its NX coordinates, then
//                      its weight when W) and index in tile order: a
//                      workgroup takes kScatT * PTS points, reads each once,
//                      orders the records by tile in LDS and writes the
//                      records of one tile as one run of consecutive lanes
//                      and addresses;
// after which one workgroup per (batch entry, tile) reads its points in order.
#pragma once

#include "internal.hpp"

namespace ddcx {
namespace binning {
namespace {

constexpr int kMaxBins = 4096;          // tiles per batch entry
constexpr int kNoBin = 0xffff;          // stored tile of a point that no tile takes
constexpr int kScatT = 1024;            // threads of the scatter workgroups
constexpr size_t kLdsBytes = 160 * 1024;

__global__ void __launch_bounds__(256) fill_int_kernel(int* __restrict__ p, int64_t n)
{
    int64_t const i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// exclusive scan of one batch entry's tile counts (nbins <= 4096, 1024 threads,
// 4 tiles each); starts gets nbins + 1 entries, cursor a copy of the first nbins
__global__ void __launch_bounds__(1024) bin_scan_kernel(int nbins, const int* __restrict__ counts,
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
        c[u] = k < nbins ? counts[l * nbins + k] : 0;
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
        if (k < nbins) {
            starts[l * (nbins + 1) + k] = run;
            cursor[l * nbins + k] = run;
        }
        run += c[u];
    }
    if (t == 1023) starts[l * (nbins + 1) + nbins] = part[1023];
}

// LDS of a scatter workgroup of kScatT * pts points with records of rw
// doubles: the records in tile order, the tile histogram, local and global
// tile offsets (nbins ints each), and per slot its tile and local index
size_t scatter_lds(int nbins, int pts, int rw)
{
    size_t const np = size_t(kScatT) * pts;
    return np * rw * sizeof(double) + size_t(3) * nbins * sizeof(int) + np * 2 * sizeof(uint16_t);
}

template <int PTS, int NX, bool W>
__global__ void __launch_bounds__(kScatT) bin_scatter_kernel(int nbins, const double* __restrict__ coords,
                                                             const double* __restrict__ w, int64_t npts, int64_t bpb,
                                                             int* __restrict__ cursor, int* __restrict__ perm,
                                                             double* __restrict__ recs,
                                                             const uint16_t* __restrict__ bins)
{
    constexpr int NP = kScatT * PTS;
    constexpr int RW = NX + (W ? 1 : 0);
    static_assert(NP <= 65536 && kMaxBins <= 65536, "slot tile and local index are 16-bit");
    extern __shared__ double srec[];
    int* const hist = reinterpret_cast<int*>(srec + RW * NP);
    int* const lofs = hist + nbins;
    int* const base = lofs + nbins;
    uint16_t* const sbin = reinterpret_cast<uint16_t*>(base + nbins);
    uint16_t* const sloc = sbin + NP;
    __shared__ int part[kScatT];
    int const tid = threadIdx.x;
    int64_t const l = blockIdx.x / bpb;
    int64_t const p0 = (blockIdx.x % bpb) * int64_t(NP);
    for (int k = tid; k < nbins; k += kScatT) hist[k] = 0;
    double rv[PTS][RW];
    int bin[PTS], rank[PTS];
#pragma unroll
    for (int u = 0; u < PTS; ++u) {
        int64_t const p = p0 + u * kScatT + tid;
        int64_t const t = l * npts + (p < npts ? p : npts - 1);
#pragma unroll
        for (int k = 0; k < NX; ++k) rv[u][k] = coords[NX * t + k];
        if constexpr (W) rv[u][NX] = w[t];
        int const b = bins[t];
        bin[u] = p < npts && b != kNoBin ? b : -1;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PTS; ++u) rank[u] = bin[u] >= 0 ? atomicAdd(&hist[bin[u]], 1) : 0;
    __syncthreads();
    // local exclusive offsets: thread t scans tiles [t*per, (t+1)*per), then
    // the kScatT chunk sums are scanned
    int const per = (nbins + kScatT - 1) / kScatT;
    int const kb = tid * per, ke = min(kb + per, nbins);
    int sum = 0;
    for (int k = kb; k < ke; ++k) sum += hist[k];
    part[tid] = sum;
    __syncthreads();
    for (int s = 1; s < kScatT; s *= 2) { // inclusive Hillis-Steele scan
        int const v = tid >= s ? part[tid - s] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    int run = part[tid] - sum;
    for (int k = kb; k < ke; ++k) {
        lofs[k] = run;
        run += hist[k];
        if (hist[k]) base[k] = atomicAdd(&cursor[l * nbins + k], hist[k]);
    }
    int const total = part[kScatT - 1];
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PTS; ++u)
        if (bin[u] >= 0) {
            int const j = lofs[bin[u]] + rank[u];
#pragma unroll
            for (int k = 0; k < RW; ++k) srec[RW * j + k] = rv[u][k];
            sbin[j] = uint16_t(bin[u]);
            sloc[j] = uint16_t(u * kScatT + tid);
        }
    __syncthreads();
    for (int j = tid; j < total; j += kScatT) {
        int const b = sbin[j];
        int64_t const k = l * npts + base[b] + (j - lofs[b]);
        perm[k] = int(p0 + sloc[j]);
#pragma unroll
        for (int c = 0; c < RW; ++c) recs[RW * k + c] = srec[RW * j + c];
    }
}

// The workspace: counts, starts, cursor (per batch entry and tile), then per
// point its index in tile order, its record (rw doubles) and its tile.
struct Work {
    int* counts;
    int* starts;
    int* cursor;
    int* perm;
    double* recs;
    uint16_t* bins;
    int64_t bytes;
};

Work carve(int nbins, int rw, int64_t npts, int64_t nbatch, void* base)
{
    auto up = [](int64_t n, int64_t size) { return (n * size + 255) / 256 * 256; };
    int64_t const a = up(nbatch * nbins, 4), b = up(nbatch * (nbins + 1), 4), c = up(nbatch * nbins, 4),
                  d = up(nbatch * npts, 4), e = up(rw * nbatch * npts, 8), f = up(nbatch * npts, 2);
    char* p = static_cast<char*>(base);
    return Work{reinterpret_cast<int*>(p),
                reinterpret_cast<int*>(p + a),
                reinterpret_cast<int*>(p + a + b),
                reinterpret_cast<int*>(p + a + b + c),
                reinterpret_cast<double*>(p + a + b + c + d),
                reinterpret_cast<uint16_t*>(p + a + b + c + d + e),
                a + b + c + d + e + f};
}

unsigned grid_blocks(int64_t n, int threads)
{
    int64_t const b = (n + threads - 1) / threads;
    if (b > int64_t(0x7fffffff)) throw Error(DDCX_ERR_INVALID, "batch too large for one launch");
    return unsigned(b);
}

// scan the counts and scatter the records (the most points per scatter
// workgroup whose records fit in LDS)
template <int NX, bool W>
void scan_and_scatter(int nbins, const double* coords, const double* w, int64_t npts, int64_t nbatch,
                      Work const& ws, hipStream_t stream)
{
    constexpr int RW = NX + (W ? 1 : 0);
    hipLaunchKernelGGL(bin_scan_kernel, dim3(unsigned(nbatch)), dim3(1024), 0, stream, nbins, ws.counts, ws.starts,
                       ws.cursor);
    auto scatter = [&]<int PTS>() {
        size_t const slds = scatter_lds(nbins, PTS, RW);
        if (slds + kScatT * sizeof(int) > kLdsBytes) return false;
        auto kern = bin_scatter_kernel<PTS, NX, W>;
        hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                      int(slds)),
                  "hipFuncSetAttribute");
        int64_t const sbpb = (npts + kScatT * PTS - 1) / (kScatT * PTS);
        hipLaunchKernelGGL(kern, dim3(grid_blocks(sbpb * nbatch, 1)), dim3(kScatT), slds, stream, nbins, coords, w,
                           npts, sbpb, ws.cursor, ws.perm, ws.recs, ws.bins);
        return true;
    };
    if (!scatter.template operator()<4>() && !scatter.template operator()<2>() && !scatter.template operator()<1>())
        throw Error(DDCX_ERR_UNSUPPORTED, "binned points: scatter does not fit in LDS");
}

} // namespace
} // namespace binning
} // namespace ddcx
