// Access-pattern ceiling for bench.py (not part of the spline library): a
// persistent copy of a (rows, lines) batch-contiguous FP64 array in tiles of
// LB lines by all rows, lanes along lines (LB * 8-byte runs per row), rows
// strided by ld * 8 bytes, non-temporal loads and stores, every element
// read and written once.  This is the pattern of the tiled spline kernels
// on batch-contiguous arrays, so its rate bounds theirs.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

template <int LB>
__global__ void __launch_bounds__(1024) tile_copy_kernel(const double* __restrict__ a, double* __restrict__ b,
                                                         int rows, int64_t lines, int64_t ld)
{
    constexpr int RG = 1024 / LB; // row groups
    constexpr int PER = 8;        // rows per thread and batch
    int const ln = threadIdx.x % LB, r0 = threadIdx.x / LB;
    int64_t const ntiles = (lines + LB - 1) / LB;
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        int64_t const l = t * LB + ln;
        if (l >= lines) continue;
        for (int rb = 0; rb < rows; rb += RG * PER) {
            double v[PER];
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                int const r = rb + r0 + RG * k;
                v[k] = r < rows ? __builtin_nontemporal_load(&a[int64_t(r) * ld + l]) : 0.0;
            }
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                int const r = rb + r0 + RG * k;
                if (r < rows) __builtin_nontemporal_store(v[k], &b[int64_t(r) * ld + l]);
            }
        }
    }
}

} // namespace

// Copies a -> b (both rows x ld doubles, the first `lines` columns) in tiles
// of lb lines (16 or 32) on `grid` workgroups; returns 0, or 1 for an
// unsupported lb or a launch error.
extern "C" int ddcx_probe_tile_copy(const double* a, double* b, int rows, int64_t lines, int64_t ld, int lb, int grid,
                                    hipStream_t stream)
{
    if (rows <= 0 || lines <= 0 || ld < lines || grid <= 0) return 1;
    if (lb == 16) hipLaunchKernelGGL(tile_copy_kernel<16>, dim3(grid), dim3(1024), 0, stream, a, b, rows, lines, ld);
    else if (lb == 32) hipLaunchKernelGGL(tile_copy_kernel<32>, dim3(grid), dim3(1024), 0, stream, a, b, rows, lines, ld);
    else return 1;
    return hipGetLastError() == hipSuccess ? 0 : 1;
}
