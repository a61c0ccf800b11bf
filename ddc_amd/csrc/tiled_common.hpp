// Device helpers shared by the register-resident build kernels
// (build_tiled_kernel.hpp, build_refine_kernel.hpp, build_vjp_tiled.hip,
// fused_build_eval.hip).
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

// The border data a workgroup reads after its sweeps, copied to LDS in its
// prologue so that the border step waits on no global load:
// - the lambda couplings (row, column, value);
// - the Q^-1 gamma blocks with each row padded to kNbMax columns (zeros past
//   nb), so that the correction is kNbMax unconditional FMAs per row (an
//   FMA with a zero entry and a zero x2 leaves the row exactly as it was).
struct BorderLds {
    int c[kLamMax];
    int r[kLamMax];
    double v[kLamMax];
};

// NT threads (NT >= kLamMax); GAMMA: also the padded Q^-1 gamma blocks
template <int NT, bool GAMMA>
__device__ __forceinline__ void stage_border(TiledView const& tv, BorderLds& bl, double* gpad, int tid)
{
    static_assert(NT >= kLamMax, "one coupling per thread");
    if (tid < tv.bl_nnz) {
        bl.c[tid] = tv.bl_c[tid];
        bl.r[tid] = tv.bl_r[tid];
        bl.v[tid] = tv.bl_v[tid];
    }
    if constexpr (GAMMA) {
        int const n = tv.ngb * tv.L * kNbMax;
        for (int idx = tid; idx < n; idx += NT) {
            int const k = idx % kNbMax;
            gpad[idx] = k < tv.nb ? tv.Gblk[(idx / kNbMax) * tv.nb + k] : 0.0;
        }
    }
}

// y -= G x2 for one segment's L x kNbMax padded block G
template <int L>
__device__ __forceinline__ void apply_gamma(double (&y)[L], const double* G, double const (&x2)[kNbMax])
{
#pragma unroll
    for (int i = 0; i < L; ++i) {
        double t = y[i];
#pragma unroll
        for (int k = 0; k < kNbMax; ++k) t = fma(-G[i * kNbMax + k], x2[k], t);
        y[i] = t;
        chunk_fence(i);
    }
}

// doubles of LDS that stage_border's padded Q^-1 gamma blocks take
inline size_t gamma_lds_doubles(TiledView const& tv) { return size_t(tv.ngb) * tv.L * kNbMax; }

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

// Interpolation-contiguous layouts (IC) stage the tile through LDS in chunks
// of kStageRows rows of all LB lines: each line's rows are read and written
// as contiguous runs, and every thread takes its L rows from LDS.
constexpr int kStageRows = 256;
constexpr int kStagePitch = kStageRows + 1; // doubles; odd pitch spreads the LDS banks
// the chunk of a K x L tiling: the whole padded line when it is shorter
// (K * L rows; the short tilings have K * L = 128 or 144, so the pitch stays odd)
constexpr int stage_rows(int kl) { return kl < kStageRows ? kl : kStageRows; }
constexpr int stage_pitch(int kl) { return stage_rows(kl) + 1; }

// The tiled kernels address a tile as a workgroup-uniform base plus 32-bit
// per-lane offsets; true when every offset of the batch fits.
inline bool offsets_fit(Layout const& l, int64_t rows, int64_t nbatch)
{
    if (l.rs < 0 || l.is < 0 || l.os < 0 || l.inner < 1) return false;
    if (nbatch > l.inner && l.os < (l.inner - 1) * l.is + 1) return false;
    int64_t const span = rows * l.rs + std::min<int64_t>(nbatch, l.inner) * l.is + l.os;
    return span < (int64_t(1) << 31);
}


// The two sweeps of the main-block solve on a K x L tiling (one thread per
// segment and line, L rows in registers): local pass with a zero incoming
// state, exchange of the outgoing states through LDS, exact incoming state
// from the transfer products, correction by the homogeneous recurrence.
// Row records: the forward sweep reads Lf(KF) at OLF and 1/d at ORDF (FDIV),
// the backward sweep Ub(KB) at OUB and 1/d at ORDB; the records of the full
// row table are [Lf(KF) | Ub(KB) | rDb | rDf (FDIV)].  The forward sweep ends
// with a barrier; the backward sweep does not.
template <int LB, int K, int L, int KF, bool FDIV, int OLF, int ORDF>
__device__ __forceinline__ void tiled_forward(double (&y)[L], const double* T, int RS, TiledView const& tv,
                                              double (*ex)[K][LB], int seg, int ln)
{
    double s[KF];
#pragma unroll
    for (int k = 0; k < KF; ++k) s[k] = 0.0;
#pragma unroll
    for (int i = 0; i < L; ++i) {
        const double* R = T + i * RS;
        double t = y[i];
#pragma unroll
        for (int k = KF; k >= 1; --k) t = fma(-s[k - 1], R[OLF + k - 1], t);
        if constexpr (FDIV) t = t * R[ORDF];
        y[i] = t;
#pragma unroll
        for (int k = KF - 1; k > 0; --k) s[k] = s[k - 1];
        s[0] = t;
        chunk_fence(i);
    }
#pragma unroll
    for (int k = 0; k < KF; ++k) ex[k][seg][ln] = s[k];
    __syncthreads();
    double in[KF];
#pragma unroll
    for (int k = 0; k < KF; ++k) in[k] = 0.0;
    for (int w = 1; w <= tv.wf; ++w) {
        int const src = seg - w;
        if (src >= 0) {
            const double* m = tv.Mf + (size_t(seg) * tv.wf + (w - 1)) * (KF * KF);
#pragma unroll
            for (int a = 0; a < KF; ++a)
#pragma unroll
                for (int b = 0; b < KF; ++b) in[a] = fma(m[a * KF + b], ex[b][src][ln], in[a]);
        }
    }
    if (seg > 0) {
        // rows += homogeneous forward recurrence started from the incoming state
#pragma unroll
        for (int i = 0; i < L; ++i) {
            const double* R = T + i * RS;
            double t = 0.0;
#pragma unroll
            for (int k = KF; k >= 1; --k) t = fma(-in[k - 1], R[OLF + k - 1], t);
            if constexpr (FDIV) t = t * R[ORDF];
            y[i] += t;
#pragma unroll
            for (int k = KF - 1; k > 0; --k) in[k] = in[k - 1];
            in[0] = t;
            chunk_fence(i);
        }
    }
    __syncthreads();
}

template <int LB, int K, int L, int KB, bool BDIVF, int OUB, int ORDB>
__device__ __forceinline__ void tiled_backward(double (&y)[L], const double* T, int RS, TiledView const& tv,
                                               double (*ex)[K][LB], int seg, int ln)
{
    double s[KB];
#pragma unroll
    for (int k = 0; k < KB; ++k) s[k] = 0.0;
#pragma unroll
    for (int i = L - 1; i >= 0; --i) {
        const double* R = T + i * RS;
        double t;
        if constexpr (BDIVF) {
            t = fma(-s[0], R[OUB], y[i] * R[ORDB]);
        } else {
            t = y[i];
#pragma unroll
            for (int k = KB; k >= 1; --k) t = fma(-s[k - 1], R[OUB + k - 1], t);
            t = t * R[ORDB];
        }
        y[i] = t;
#pragma unroll
        for (int k = KB - 1; k > 0; --k) s[k] = s[k - 1];
        s[0] = t;
        chunk_fence(i);
    }
#pragma unroll
    for (int k = 0; k < KB; ++k) ex[k][seg][ln] = s[k];
    __syncthreads();
    double in[KB];
#pragma unroll
    for (int k = 0; k < KB; ++k) in[k] = 0.0;
    for (int w = 1; w <= tv.wb; ++w) {
        int const src = seg + w;
        if (src < K) {
            const double* m = tv.Mb + (size_t(seg) * tv.wb + (w - 1)) * (KB * KB);
#pragma unroll
            for (int a = 0; a < KB; ++a)
#pragma unroll
                for (int b = 0; b < KB; ++b) in[a] = fma(m[a * KB + b], ex[b][src][ln], in[a]);
        }
    }
    if (seg + 1 < K) {
        // rows += homogeneous backward recurrence started from the incoming state
#pragma unroll
        for (int i = L - 1; i >= 0; --i) {
            const double* R = T + i * RS;
            double t;
            if constexpr (BDIVF) {
                t = -in[0] * R[OUB];
            } else {
                t = 0.0;
#pragma unroll
                for (int k = KB; k >= 1; --k) t = fma(-in[k - 1], R[OUB + k - 1], t);
                t = t * R[ORDB];
            }
            y[i] += t;
#pragma unroll
            for (int k = KB - 1; k > 0; --k) in[k] = in[k - 1];
            in[0] = t;
            chunk_fence(i);
        }
    }
}

template <int LB, int K, int L, int KF, int KB, bool FDIV, bool BDIVF, int KE>
__device__ __forceinline__ void tiled_sweeps(double (&y)[L], const double* T, int RS, TiledView const& tv,
                                             double (*ex)[K][LB], int seg, int ln)
{
    tiled_forward<LB, K, L, KF, FDIV, 0, KF + KB + 1>(y, T, RS, tv, ex, seg, ln);
    tiled_backward<LB, K, L, KB, BDIVF, KF, KF + KB>(y, T, RS, tv, ex, seg, ln);
}

// Split staging of the row table, for tables that do not fit LDS whole:
// the fields one sweep reads, copied from the full records (stride RS) into
// compact records in LDS ([Lf(KF) | rDf] for the forward sweep, [Ub(KB) |
// rDb] for the backward one).  The caller synchronises.
template <int NT, int KF, int KB, bool FDIV, bool FORWARD>
__device__ __forceinline__ void stage_half(const double* __restrict__ rt, int RS, int rows, double* dst, int tid)
{
    constexpr int W = FORWARD ? KF + (FDIV ? 1 : 0) : KB + 1;
    for (int e = tid; e < rows * W; e += NT) {
        int const r = e / W, f = e % W;
        int src;
        if constexpr (FORWARD) src = f < KF ? f : KF + KB + 1;
        else src = f < KB ? KF + f : KF + KB;
        dst[e] = rt[size_t(r) * RS + src];
    }
}
} // namespace tiled
} // namespace ddcx
