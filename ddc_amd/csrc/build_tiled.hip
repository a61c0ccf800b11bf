// Register-resident tiled batched spline build (the hot path of
// SplineBuilder::operator(), spline_builder.hpp:896-1101).
//
// A workgroup of LB*K threads owns LB batch lines; thread (segment s, line l)
// holds rows [s*L, s*L + L) of line l's main-block right-hand side in
// registers.  Each line is therefore read from HBM once and written once:
//   1. load the L rows (lanes of a wave cover LB consecutive lines, so each
//      load instruction reads 64/LB contiguous runs of LB*8 bytes in the
//      batch-contiguous layout);
//   2. forward sweep, local pass with a zero incoming state -> outgoing
//      state, exchanged through LDS; the exact incoming state of every
//      segment is a short sum of precomputed transfer products; second pass
//      with the true state (the same row operations, in the same order, as
//      the sequential LAPACK sweep);
//   3. the same for the backward sweep;
//   4. 2x2-block border (splines_linear_problem_2x2_blocks.cpp:275-298):
//      lambda couplings gathered through LDS, the small Schur system solved
//      by one thread per line, the Q^-1 gamma correction applied in
//      registers;
//   5. store, including the periodic duplicate rows (spline_builder.hpp:
//      1074-1100).
#include "internal.hpp"

namespace ddcx {

namespace {

constexpr int kNbMax = 8;
constexpr int kLamMax = 64;

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

template <int LB, int K, int L, int KF, int KB, bool FDIV, bool BDIVF>
__global__ void __launch_bounds__(LB* K) build_tiled_kernel(TiledView tv, BuildView bv,
                                                           const double* __restrict__ vals, Layout vl,
                                                           const double* __restrict__ dl,
                                                           const double* __restrict__ dr, Layout dlay,
                                                           double* __restrict__ coef, Layout cl, int64_t nbatch)
{
    constexpr int KE = (KF > KB ? KF : KB) > 0 ? (KF > KB ? KF : KB) : 1;
    constexpr int SF = KF > 0 ? KF : 1;
    constexpr int SB = KB > 0 ? KB : 1;
    __shared__ double ex[KE][K][LB];
    __shared__ double lamx[kLamMax][LB];
    __shared__ double x2s[kNbMax][LB];

    int const tid = threadIdx.x;
    int const ln = tid % LB;
    int const seg = tid / LB;
    int64_t const line = int64_t(blockIdx.x) * LB + ln;
    bool const active = line < nbatch;
    int64_t const lc = active ? line : 0;
    int64_t const vb = lay(vl, 0, lc);
    int64_t const cb = lay(cl, 0, lc);
    int const r0 = seg * L;
    int const nq = tv.nq;

    double y[L];
#pragma unroll
    for (int i = 0; i < L; ++i) {
        int const row = r0 + i;
        y[i] = (active && row < nq) ? vals[vb + int64_t(tv.q0 + row) * vl.rs] : 0.0;
    }

    // ---------------- forward sweep ----------------
    if constexpr (KF > 0) {
        double s[SF];
#pragma unroll
        for (int k = 0; k < SF; ++k) s[k] = 0.0;
#pragma unroll
        for (int i = 0; i < L; ++i) {
            int const row = r0 + i;
            double t = y[i];
#pragma unroll
            for (int k = KF; k >= 1; --k) t = t - s[k - 1] * tv.Lf[row * KF + (k - 1)];
            if constexpr (FDIV) t = t / tv.Df[row];
#pragma unroll
            for (int k = SF - 1; k > 0; --k) s[k] = s[k - 1];
            s[0] = t;
        }
#pragma unroll
        for (int k = 0; k < KF; ++k) ex[k][seg][ln] = s[k];
        __syncthreads();
        double in[SF];
#pragma unroll
        for (int k = 0; k < SF; ++k) in[k] = 0.0;
        for (int w = 1; w <= tv.wf; ++w) {
            int const src = seg - w;
            if (src >= 0) {
                const double* m = tv.Mf + (size_t(seg) * tv.wf + (w - 1)) * (KF * KF);
#pragma unroll
                for (int a = 0; a < KF; ++a)
#pragma unroll
                    for (int b = 0; b < KF; ++b) in[a] += m[a * KF + b] * ex[b][src][ln];
            }
        }
#pragma unroll
        for (int k = 0; k < SF; ++k) s[k] = in[k];
#pragma unroll
        for (int i = 0; i < L; ++i) {
            int const row = r0 + i;
            double t = y[i];
#pragma unroll
            for (int k = KF; k >= 1; --k) t = t - s[k - 1] * tv.Lf[row * KF + (k - 1)];
            if constexpr (FDIV) t = t / tv.Df[row];
            y[i] = t;
#pragma unroll
            for (int k = SF - 1; k > 0; --k) s[k] = s[k - 1];
            s[0] = t;
        }
        __syncthreads();
    } else if constexpr (FDIV) {
#pragma unroll
        for (int i = 0; i < L; ++i) y[i] = y[i] / tv.Df[r0 + i];
    }

    // ---------------- backward sweep ----------------
    auto bwd_row = [&](int row, double yi, const double* s) {
        if constexpr (BDIVF) {
            return yi / tv.Db[row] - s[0] * tv.Ub[row * SB];
        } else {
            double t = yi;
#pragma unroll
            for (int k = KB; k >= 1; --k) t = t - s[k - 1] * tv.Ub[row * KB + (k - 1)];
            return t / tv.Db[row];
        }
    };
    if constexpr (KB > 0) {
        double s[SB];
#pragma unroll
        for (int k = 0; k < SB; ++k) s[k] = 0.0;
#pragma unroll
        for (int i = L - 1; i >= 0; --i) {
            double const t = bwd_row(r0 + i, y[i], s);
#pragma unroll
            for (int k = SB - 1; k > 0; --k) s[k] = s[k - 1];
            s[0] = t;
        }
#pragma unroll
        for (int k = 0; k < KB; ++k) ex[k][seg][ln] = s[k];
        __syncthreads();
        double in[SB];
#pragma unroll
        for (int k = 0; k < SB; ++k) in[k] = 0.0;
        for (int w = 1; w <= tv.wb; ++w) {
            int const src = seg + w;
            if (src < K) {
                const double* m = tv.Mb + (size_t(seg) * tv.wb + (w - 1)) * (KB * KB);
#pragma unroll
                for (int a = 0; a < KB; ++a)
#pragma unroll
                    for (int b = 0; b < KB; ++b) in[a] += m[a * KB + b] * ex[b][src][ln];
            }
        }
#pragma unroll
        for (int k = 0; k < SB; ++k) s[k] = in[k];
#pragma unroll
        for (int i = L - 1; i >= 0; --i) {
            double const t = bwd_row(r0 + i, y[i], s);
            y[i] = t;
#pragma unroll
            for (int k = SB - 1; k > 0; --k) s[k] = s[k - 1];
            s[0] = t;
        }
    } else {
#pragma unroll
        for (int i = 0; i < L; ++i) y[i] = y[i] / tv.Db[r0 + i];
    }

    // ---------------- border (2x2 blocks) ----------------
    int const nb = tv.nb;
    int const off = bv.offset;
    int const nbas = bv.nbasis;
    if (nb > 0) {
        for (int z = 0; z < tv.bl_nnz; ++z) {
            int const c = tv.bl_c[z];
            if (c >= r0 && c < r0 + L) {
                double v = 0.0;
#pragma unroll
                for (int i = 0; i < L; ++i)
                    if (r0 + i == c) v = y[i];
                lamx[z][ln] = v;
            }
        }
        __syncthreads();
        if (seg == 0) {
            double b2[kNbMax];
            int64_t const db = lay(dlay, 0, lc);
#pragma unroll
            for (int k = 0; k < kNbMax; ++k)
                b2[k] = (active && k < nb) ? rhs_src(bv, vals, vl, vb, dl, dr, dlay, db, k < tv.top ? k : nq + k)
                                           : 0.0;
            for (int z = 0; z < tv.bl_nnz; ++z) {
                int const r = tv.bl_r[z];
                double const v = tv.bl_v[z] * lamx[z][ln];
#pragma unroll
                for (int k = 0; k < kNbMax; ++k)
                    if (k == r) b2[k] -= v;
            }
            getrs_small(nb, tv.s_lu, tv.s_ipiv, b2);
#pragma unroll
            for (int k = 0; k < kNbMax; ++k) {
                if (k < nb) {
                    x2s[k][ln] = b2[k];
                    if (active) {
                        int const prow = k < tv.top ? k : nq + k;
                        coef[cb + int64_t(prow + off) * cl.rs] = b2[k];
                        if (bv.periodic) {
                            int const j0 = prow + off - nbas;
                            if (j0 >= 0) coef[cb + int64_t(j0) * cl.rs] = b2[k];
                            int const j2 = prow + off + nbas;
                            if (j2 < nbas + bv.degree) coef[cb + int64_t(j2) * cl.rs] = b2[k];
                        }
                    }
                }
            }
        }
        __syncthreads();
        double x2[kNbMax];
#pragma unroll
        for (int k = 0; k < kNbMax; ++k) x2[k] = k < nb ? x2s[k][ln] : 0.0;
#pragma unroll
        for (int i = 0; i < L; ++i) {
            int const row = r0 + i;
            double t = y[i];
#pragma unroll
            for (int k = 0; k < kNbMax; ++k)
                if (k < nb) t = t - tv.G[row * nb + k] * x2[k];
            y[i] = t;
        }
    }

    // ---------------- store ----------------
    if (active) {
#pragma unroll
        for (int i = 0; i < L; ++i) {
            int const row = r0 + i;
            if (row < nq) {
                int const prow = tv.top + row;
                coef[cb + int64_t(prow + off) * cl.rs] = y[i];
                if (bv.periodic) {
                    int const j0 = prow + off - nbas;
                    if (j0 >= 0) coef[cb + int64_t(j0) * cl.rs] = y[i];
                    int const j2 = prow + off + nbas;
                    if (j2 < nbas + bv.degree) coef[cb + int64_t(j2) * cl.rs] = y[i];
                }
            }
        }
    }
}

template <int L, int KF, int KB, bool FDIV, bool BDIVF>
void launch_cfg(BuilderPlan const& p, double* coef, Layout cl, const double* vals, Layout vl, const double* dl,
                const double* dr, Layout dlay, int64_t nbatch, hipStream_t stream)
{
    constexpr int LB = 16, K = 64;
    int64_t const blocks = (nbatch + LB - 1) / LB;
    hipLaunchKernelGGL((build_tiled_kernel<LB, K, L, KF, KB, FDIV, BDIVF>), dim3(unsigned(blocks)), dim3(LB * K), 0,
                       stream, p.tview, p.view, vals, vl, dl, dr, dlay, coef, cl, nbatch);
    hip_check(hipGetLastError(), "build_tiled_kernel launch");
}

template <int L>
bool launch_L(BuilderPlan const& p, double* coef, Layout cl, const double* vals, Layout vl, const double* dl,
              const double* dr, Layout dlay, int64_t nbatch, hipStream_t stream)
{
    TiledHost const& t = p.tiled;
    auto go = [&]<int KF, int KB, bool FD, bool BD>() {
        launch_cfg<L, KF, KB, FD, BD>(p, coef, cl, vals, vl, dl, dr, dlay, nbatch, stream);
        return true;
    };
    if (t.bwd_divfirst) {
        if (t.kf == 1 && t.kb == 1 && !t.fwd_div) return go.template operator()<1, 1, false, true>();
        return false;
    }
    if (t.fwd_div) { // PDSBand
        if (t.kf == 2 && t.kb == 2) return go.template operator()<2, 2, true, false>();
        if (t.kf == 3 && t.kb == 3) return go.template operator()<3, 3, true, false>();
        return false;
    }
    // unpivoted band
    if (t.kf == 1 && t.kb == 1) return go.template operator()<1, 1, false, false>();
    if (t.kf == 2 && t.kb == 2) return go.template operator()<2, 2, false, false>();
    if (t.kf == 3 && t.kb == 3) return go.template operator()<3, 3, false, false>();
    if (t.kf == 4 && t.kb == 4) return go.template operator()<4, 4, false, false>();
    return false;
}

} // namespace

bool launch_build_tiled(BuilderPlan const& p, double* coef, Layout cl, const double* vals, Layout vl,
                        const double* dl, const double* dr, Layout dlay, int64_t nbatch, hipStream_t stream)
{
    if (!p.tiled.ok) return false;
    if (p.tiled.L == 16) return launch_L<16>(p, coef, cl, vals, vl, dl, dr, dlay, nbatch, stream);
    if (p.tiled.L == 32) return launch_L<32>(p, coef, cl, vals, vl, dl, dr, dlay, nbatch, stream);
    return false;
}

} // namespace ddcx
