// Device helpers shared by the register-resident build kernels
// (build_tiled.hip, build_refine.hip).
#pragma once

#include <algorithm>

#include "internal.hpp"

namespace ddcx {
namespace tiled {

constexpr int kNbMax = 8;   // border (Schur) size limit
constexpr int kLamMax = 64; // lambda couplings limit

// Keeps the scheduler from hoisting the table loads of a whole segment at
// once (which would not fit in registers next to the segment's rows): a
// scheduling barrier after every M rows of an unrolled row loop.
template <int M = 8>
__device__ __forceinline__ void chunk_fence(int i)
{
    if (i % M == M - 1) __builtin_amdgcn_sched_barrier(0);
}

__device__ __forceinline__ int64_t lay(Layout const& L, int64_t r, int64_t l)
{
    return r * L.rs + (l % L.inner) * L.is + (l / L.inner) * L.os;
}

// physical right-hand-side row i of one line (spline_builder.hpp:949-1033)
__device__ __forceinline__ double rhs_src(BuildView const& bv, const double* vals, Layout const& vl, int64_t vb,
                                          const double* dl, const double* dr, Layout const& dlay, int64_t db, int i)
{
    if (bv.periodic) return vals[vb + int64_t(i) * vl.rs];
    if (i < bv.nbe_l) return bv.bcl == DDCX_HERMITE ? dl[db + int64_t(i) * dlay.rs] * bv.dxpow[i + bv.odd] : 0.0;
    int const nv_end = bv.nbasis - bv.nbe_r;
    if (i < nv_end) return vals[vb + int64_t(i - bv.nbe_l) * vl.rs];
    return bv.bcu == DDCX_HERMITE ? dr[db + int64_t(i - nv_end) * dlay.rs] * bv.dxpow[i - nv_end + bv.odd] : 0.0;
}

// dgetrs on an nb x nb LU (nb <= kNbMax) held in registers
__device__ __forceinline__ void getrs_small(int nb, const double* a, const int* ipiv, double* b)
{
#pragma unroll
    for (int j = 0; j < kNbMax; ++j) {
        if (j < nb) {
            int const p = ipiv[j];
            double bp = 0.0;
#pragma unroll
            for (int k = 0; k < kNbMax; ++k)
                if (k == p) bp = b[k];
            double const bj = b[j];
#pragma unroll
            for (int k = 0; k < kNbMax; ++k)
                if (k == p) b[k] = bj;
            b[j] = bp;
        }
    }
#pragma unroll
    for (int j = 0; j < kNbMax; ++j)
#pragma unroll
        for (int i = j + 1; i < kNbMax; ++i)
            if (i < nb) b[i] = b[i] - b[j] * a[i * nb + j];
#pragma unroll
    for (int j = kNbMax - 1; j >= 0; --j) {
        if (j < nb) {
            b[j] = b[j] / a[j * nb + j];
#pragma unroll
            for (int i = 0; i < j; ++i) b[i] = b[i] - b[j] * a[i * nb + j];
        }
    }
}

// The tiled kernels address a tile as a workgroup-uniform base plus 32-bit
// per-lane offsets; true when every offset of the batch fits.
inline bool offsets_fit(Layout const& l, int64_t rows, int64_t nbatch)
{
    if (l.rs < 0 || l.is < 0 || l.os < 0 || l.inner < 1) return false;
    if (nbatch > l.inner && l.os < (l.inner - 1) * l.is + 1) return false;
    int64_t const span = rows * l.rs + std::min<int64_t>(nbatch, l.inner) * l.is + l.os;
    return span < (int64_t(1) << 31);
}

} // namespace tiled
} // namespace ddcx
