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
//      segment is a short sum of precomputed transfer products, and the
//      segment's rows are corrected by the homogeneous recurrence of the
//      sweep started from that state;
//   3. the same for the backward sweep;
//   4. 2x2-block border (splines_linear_problem_2x2_blocks.cpp:275-298):
//      lambda couplings gathered through LDS, the small Schur system solved
//      by one thread per line, the Q^-1 gamma correction applied in
//      registers;
//   5. store, including the periodic duplicate rows (spline_builder.hpp:
//      1074-1100).
#pragma once

#include <algorithm>

#include "tiled_common.hpp"

namespace ddcx {
namespace tiled_build {

using namespace tiled;

// TMODE: 0 row table read from global memory (L2), 1 staged in LDS whole,
// 2 staged in LDS one sweep's fields at a time (tables too large for 1)
template <int LB, int K, int L, int KF, int KB, bool FDIV, bool BDIVF, int TMODE, bool IC>
__global__ void __launch_bounds__(LB* K) build_tiled_kernel(TiledView tv, BuildView bv,
                                                           const double* __restrict__ vals, Layout vl,
                                                           const double* __restrict__ dl,
                                                           const double* __restrict__ dr, Layout dlay,
                                                           double* __restrict__ coef, Layout cl, int64_t nbatch)
{
    constexpr int KE = KF > KB ? KF : KB;
    constexpr int NT = LB * K;
    constexpr int SR = stage_rows(K * L), SP = stage_pitch(K * L); // staging chunk rows, pitch (IC)
    constexpr bool TLDS = TMODE != 0;
    constexpr int RSF = KF + (FDIV ? 1 : 0), RSB = KB + 1; // split records
    __shared__ double ex[KE][K][LB];
    __shared__ double lamx[kLamMax][LB];
    __shared__ double x2s[kNbMax][LB];
    __shared__ BorderLds bl;
    extern __shared__ double rt_lds[];

    int const tid = threadIdx.x;
    int const ln = tid % LB;
    int const seg = tid / LB;
    int64_t const line0 = int64_t(blockIdx.x) * LB;
    int64_t const line = line0 + ln;
    bool const active = line < nbatch;
    int64_t const lc = active ? line : line0;
    int64_t const vb = lay(vl, 0, lc);
    int64_t const cb = lay(cl, 0, lc);
    int const r0 = seg * L;
    int const nq = tv.nq;
    constexpr int RS = KF + KB + 1 + (FDIV ? 1 : 0); // row record length, fixed by the band shape
    // Workgroup-uniform bases plus 32-bit per-lane offsets (the host checks
    // that every offset of a tile fits), so that each load/store needs one
    // address register instead of two.
    const double* __restrict__ vwg = vals + lay(vl, 0, line0);
    double* __restrict__ cwg = coef + lay(cl, 0, line0);
    uint32_t const vlo = uint32_t(vb - lay(vl, 0, line0)) + uint32_t(tv.q0 + r0) * uint32_t(vl.rs);
    uint32_t const clo = uint32_t(cb - lay(cl, 0, line0));
    uint32_t const vrs = uint32_t(vl.rs), crs = uint32_t(cl.rs);

    // dynamic LDS: the row table and the padded Q^-1 gamma blocks (TLDS),
    // then the staging area, LB x stage_pitch(K * L) doubles, and the per-line
    // offsets of vals and coef (IC)
    int const tab = TMODE == 2 ? K * L * (RSF > RSB ? RSF : RSB) : K * L * RS; // doubles of the table region
    double* gpad = rt_lds + tab;
    double* stg = rt_lds + (TLDS ? tab + tv.ngb * L * kNbMax : 0);
    uint32_t* loff = reinterpret_cast<uint32_t*>(stg + LB * SP);
    int const nl = int(min<int64_t>(LB, nbatch - line0));

    double y[L];
    if constexpr (IC) {
        constexpr int PER = LB * SR / NT; // staged elements per thread and chunk
        if (tid < LB) {
            int64_t const l = line0 + min(tid, nl - 1);
            loff[tid] = uint32_t(lay(vl, 0, l) - lay(vl, 0, line0));
            loff[LB + tid] = uint32_t(lay(cl, 0, l) - lay(cl, 0, line0));
        }
        __syncthreads();
        for (int rbase = 0; rbase < K * L; rbase += SR) {
            double tmp[PER];
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                int const e = tid + k * NT;
                int const l = e / SR, u = rbase + e % SR;
                tmp[k] = (l < nl && u < nq) ? __builtin_nontemporal_load(&vwg[loff[l] + uint32_t(tv.q0 + u)]) : 0.0;
            }
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                int const e = tid + k * NT;
                stg[(e / SR) * SP + e % SR] = tmp[k];
            }
            __syncthreads();
            if (r0 >= rbase && r0 < rbase + SR) {
#pragma unroll
                for (int i = 0; i < L; ++i) y[i] = stg[ln * SP + (r0 - rbase) + i];
            }
            __syncthreads();
        }
    } else if (r0 + L <= nq) {
        // values and coefficients are touched once: non-temporal loads and
        // stores keep them from displacing what the next kernel reads
#pragma unroll
        for (int i = 0; i < L; ++i) y[i] = __builtin_nontemporal_load(&vwg[vlo + uint32_t(i) * vrs]);
    } else {
        // last segments: rows at or past nq read row nq - 1 (in bounds) and are zeroed
        uint32_t const vl0 = vlo - uint32_t(r0) * vrs; // row q0 of this line
#pragma unroll
        for (int i = 0; i < L; ++i) {
            double const v = __builtin_nontemporal_load(&vwg[vl0 + uint32_t(min(r0 + i, nq - 1)) * vrs]);
            y[i] = r0 + i < nq ? v : 0.0;
        }
    }

    // the border rows' right-hand side does not depend on the sweeps: segment
    // 0 loads it now, with the tile, and parks it in x2s (which it overwrites
    // with the border solution later), so the border step does not wait on
    // memory while the other waves wait at its barrier
    if (tv.nb > 0 && seg == 0) {
        int64_t const db = lay(dlay, 0, lc);
#pragma unroll
        for (int k = 0; k < kNbMax; ++k)
            if (k < tv.nb)
                x2s[k][ln] = active ? rhs_src(bv, vals, vl, vb, dl, dr, dlay, db, k < tv.top ? k : tv.nq + k) : 0.0;
    }

    // the segment's Q^-1 gamma block (-1: none), read before it is needed
    int const gblk = tv.nb > 0 ? tv.gseg[seg] : -1;

    // per-row table and border data: staged in LDS when they fit (one copy
    // per workgroup)
    if constexpr (TMODE == 2) {
        // the forward sweep's fields, then (after its closing barrier) the
        // backward sweep's, in the same LDS region
        stage_half<NT, KF, KB, FDIV, true>(tv.rt, RS, K * L, rt_lds, tid);
        stage_border<NT, true>(tv, bl, gpad, tid);
        __syncthreads();
        tiled_forward<LB, K, L, KF, FDIV, 0, KF>(y, rt_lds + r0 * RSF, RSF, tv, ex, seg, ln);
        stage_half<NT, KF, KB, FDIV, false>(tv.rt, RS, K * L, rt_lds, tid);
        __syncthreads();
        tiled_backward<LB, K, L, KB, BDIVF, 0, KB>(y, rt_lds + r0 * RSB, RSB, tv, ex, seg, ln);
    } else {
        const double* T;
        if constexpr (TLDS) {
            // a few double2 per thread (RS <= 10 doubles per row, K * L <= 2 * NT rows)
            int const n2 = (K * L * RS) / 2;
            const double2* src = reinterpret_cast<const double2*>(tv.rt);
            double2* dst = reinterpret_cast<double2*>(rt_lds);
            for (int idx = tid; idx < n2; idx += NT) dst[idx] = src[idx];
            stage_border<NT, true>(tv, bl, gpad, tid);
            __syncthreads();
            T = rt_lds + r0 * RS;
        } else {
            stage_border<NT, false>(tv, bl, gpad, tid); // couplings only
            __syncthreads();
            T = tv.rt + size_t(r0) * RS;
        }
        tiled_sweeps<LB, K, L, KF, KB, FDIV, BDIVF, KE>(y, T, RS, tv, ex, seg, ln);
    }

    // ---------------- border (2x2 blocks) ----------------
    int const nb = tv.nb;
    int const off = bv.offset;
    int const nbas = bv.nbasis;
    if (nb > 0) {
        for (int z = 0; z < tv.bl_nnz; ++z) {
            int const c = bl.c[z];
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
#pragma unroll
            for (int k = 0; k < kNbMax; ++k) b2[k] = k < nb ? x2s[k][ln] : 0.0;
            for (int z = 0; z < tv.bl_nnz; ++z) {
                int const r = bl.r[z];
                double const v = bl.v[z] * lamx[z][ln];
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
        if (gblk >= 0) { // Q^-1 gamma is nonzero only in the segments near the block ends
            double x2[kNbMax];
#pragma unroll
            for (int k = 0; k < kNbMax; ++k) x2[k] = k < nb ? x2s[k][ln] : 0.0;
            if constexpr (TLDS) {
                apply_gamma<L>(y, gpad + gblk * L * kNbMax, x2);
            } else {
                const double* Gb = tv.Gblk + size_t(gblk) * L * nb;
#pragma unroll
                for (int i = 0; i < L; ++i) {
                    double t = y[i];
#pragma unroll
                    for (int k = 0; k < kNbMax; ++k)
                        if (k < nb) t = fma(-Gb[i * nb + k], x2[k], t);
                    y[i] = t;
                    chunk_fence(i);
                }
            }
        }
    }

    // ---------------- store ----------------
    if constexpr (IC) {
        constexpr int PER = LB * SR / NT;
        for (int rbase = 0; rbase < K * L; rbase += SR) {
            if (r0 >= rbase && r0 < rbase + SR) {
#pragma unroll
                for (int i = 0; i < L; ++i) stg[ln * SP + (r0 - rbase) + i] = y[i];
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                int const e = tid + k * NT;
                int const l = e / SR, u = rbase + e % SR;
                if (l < nl && u < nq)
                    __builtin_nontemporal_store(stg[l * SP + e % SR],
                                                &cwg[loff[LB + l] + uint32_t(tv.top + u + off)]);
            }
            __syncthreads();
        }
    }
    if (active) {
        uint32_t const c0 = clo + uint32_t(tv.top + r0 + off) * crs;
        if constexpr (IC) {
            // rows stored above
        } else if (r0 + L <= nq) {
#pragma unroll
            for (int i = 0; i < L; ++i) __builtin_nontemporal_store(y[i], &cwg[c0 + uint32_t(i) * crs]);
        } else {
#pragma unroll
            for (int i = 0; i < L; ++i)
                if (r0 + i < nq) __builtin_nontemporal_store(y[i], &cwg[c0 + uint32_t(i) * crs]);
        }
        if (bv.periodic && (r0 < bv.degree || r0 + L > nbas - off)) {
#pragma unroll
            for (int i = 0; i < L; ++i) {
                int const prow = tv.top + r0 + i;
                if (prow < nq) {
                    int const j0 = prow + off - nbas;
                    if (j0 >= 0) cwg[clo + uint32_t(j0) * crs] = y[i];
                    int const j2 = prow + off + nbas;
                    if (j2 < nbas + bv.degree) cwg[clo + uint32_t(j2) * crs] = y[i];
                }
            }
        }
    }
}

template <int K, int L, int KF, int KB, bool FDIV, bool BDIVF, int TMODE, bool IC, int NT>
void launch_one(BuilderPlan const& p, size_t lds, double* coef, Layout cl, const double* vals, Layout vl,
                const double* dl, const double* dr, Layout dlay, int64_t nbatch, hipStream_t stream)
{
    constexpr int LB = NT / K;
    auto kern = build_tiled_kernel<LB, K, L, KF, KB, FDIV, BDIVF, TMODE, IC>;
    if (lds > 0)
        hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                      int(lds)),
                  "hipFuncSetAttribute");
    int64_t const blocks = (nbatch + LB - 1) / LB;
    hipLaunchKernelGGL(kern, dim3(unsigned(blocks)), dim3(NT), lds, stream, p.tview, p.view, vals, vl, dl, dr, dlay,
                       coef, cl, nbatch);
    hip_check(hipGetLastError(), "build_tiled_kernel launch");
}

template <int K, int L, int KF, int KB, bool FDIV, bool BDIVF, int NT>
void launch_cfg(BuilderPlan const& p, double* coef, Layout cl, const double* vals, Layout vl, const double* dl,
                const double* dr, Layout dlay, int64_t nbatch, hipStream_t stream)
{
    if (p.tiled.rs != KF + KB + 1 + (FDIV ? 1 : 0))
        throw Error(DDCX_ERR_INVALID, "row table record length does not match the band shape");
    constexpr int LB = NT / K;
    constexpr int KE = KF > KB ? KF : KB;
    constexpr size_t static_lds =
        size_t(KE * K * LB + kLamMax * LB + kNbMax * LB) * sizeof(double) + sizeof(BorderLds);
    bool const ic = vl.rs == 1 && cl.rs == 1;
    size_t const stage = ic ? size_t(LB) * stage_pitch(K * L) * sizeof(double) + 2 * LB * sizeof(uint32_t) : 0;
    constexpr int RSH = (KF + (FDIV ? 1 : 0)) > (KB + 1) ? (KF + (FDIV ? 1 : 0)) : (KB + 1);
    size_t const gam = gamma_lds_doubles(p.tview) * sizeof(double);
    size_t const table = size_t(K) * L * p.tiled.rs * sizeof(double) + gam;
    size_t const half = size_t(K) * L * RSH * sizeof(double) + gam;
    size_t constexpr budget = 160 * 1024;
    int const mode = (static_lds + table + stage <= budget && p.tiled.rs <= 10) ? 1
                     : (static_lds + half + stage <= budget)                   ? 2
                                                                               : 0;
    size_t const lds = (mode == 1 ? table : mode == 2 ? half : 0) + stage;
    auto go = [&]<int TM, bool I>() {
        launch_one<K, L, KF, KB, FDIV, BDIVF, TM, I, NT>(p, lds, coef, cl, vals, vl, dl, dr, dlay, nbatch, stream);
    };
    auto pick = [&]<bool I>() {
        if (mode == 1) go.template operator()<1, I>();
        else if (mode == 2) go.template operator()<2, I>();
        else go.template operator()<0, I>();
    };
    if (ic) pick.template operator()<true>();
    else pick.template operator()<false>();
}

// NT threads per workgroup: NT / K lines of K segments of L rows
template <int K, int L, int NT = 1024>
bool launch_K(BuilderPlan const& p, double* coef, Layout cl, const double* vals, Layout vl, const double* dl,
              const double* dr, Layout dlay, int64_t nbatch, hipStream_t stream)
{
    TiledHost const& t = p.tiled;
    auto go = [&]<int KF, int KB, bool FD, bool BD>() {
        launch_cfg<K, L, KF, KB, FD, BD, NT>(p, coef, cl, vals, vl, dl, dr, dlay, nbatch, stream);
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

} // namespace tiled_build
} // namespace ddcx
