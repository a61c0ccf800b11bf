// Mixed-precision batched spline build: FP32 sweeps with FP64 iterative
// refinement, over the same block solve as the FP64 path
// (SplineBuilder::operator(), spline_builder.hpp:896-1101;
// splines_linear_problem_2x2_blocks.cpp:275-298 for the border steps).
//
// The tiling is the one of build_tiled_kernel.hpp (thread = L rows of one line, LB
// lines per workgroup) with L = 16, so that a thread keeps its rows of the
// line's FP64 right-hand side b and FP64 solution x in registers next to the
// FP32 work rows d:
//   x = 0, r = b
//   repeat steps + 1 times (k = 0, 1, ...):
//     sc = 2^(k*shift - e), e the binary exponent of max |b| over the line
//     d  = S(sc r)   FP32 forward/backward sweeps with segment carries, then
//                    the border and Schur steps in FP64
//     x += d / sc
//     r  = b - A x   FP64: band rows of the assembled main block (a halo of
//                    kq rows each side through LDS) and the border terms
// so that HBM sees b once and x once, as in the FP64 kernel.  shift is the
// binary order of the contraction measured on the host, so the scaled
// residuals stay near magnitude 1 in every step; the residual is formed
// directly in the FP32 work rows.
#pragma once

#include <climits>

#include "tiled_common.hpp"

namespace ddcx {
namespace refine {

using namespace tiled;

constexpr int kNoExp = INT_MIN / 2;

template <int LB, int K, int L, int KF, int KB, bool FDIV, bool BDIVF, bool TLDS, int STEPS>
__global__ void __launch_bounds__(LB* K) build_refine_kernel(RefineView rv, BuildView bv,
                                                            const double* __restrict__ vals, Layout vl,
                                                            const double* __restrict__ dl,
                                                            const double* __restrict__ dr, Layout dlay,
                                                            double* __restrict__ coef, Layout cl, int64_t nbatch)
{
    constexpr int KQ = KF; // half bandwidth of the assembled main block
    constexpr int KE = KF > KB ? KF : KB;
    constexpr int KX = 2 * KQ > KE ? 2 * KQ : KE;
    constexpr int NT = LB * K;
    constexpr int W = 2 * KQ + 1;
    // row-record field offsets: [Lf(KF) | Ub(KB) | rDb | rDf (PDS band only)]
    constexpr int oLf = 0, oUb = KF, orDb = KF + KB, orDf = KF + KB + 1;
    __shared__ double xch[KX][K][LB]; // sweep states / solution halo
    __shared__ double lamx[kLamMax][LB];
    __shared__ double b2s[kNbMax][LB]; // border right-hand side
    __shared__ double r2s[kNbMax][LB]; // border residual
    __shared__ double e2s[kNbMax][LB]; // border correction of the current solve (scaled)
    __shared__ double x2s[kNbMax][LB]; // border solution
    __shared__ int lexp[LB];
    extern __shared__ float rt32_lds[];

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
    int const nq = rv.nq;
    int const nb = rv.nb;
    constexpr int RS = KF + KB + 1 + (FDIV ? 1 : 0); // row record length, fixed by the band shape
    const double* __restrict__ vwg = vals + lay(vl, 0, line0);
    double* __restrict__ cwg = coef + lay(cl, 0, line0);
    uint32_t const vlo = uint32_t(vb - lay(vl, 0, line0)) + uint32_t(rv.q0 + r0) * uint32_t(vl.rs);
    uint32_t const clo = uint32_t(cb - lay(cl, 0, line0));
    uint32_t const vrs = uint32_t(vl.rs), crs = uint32_t(cl.rs);

    double bq[L];
    if (r0 + L <= nq) {
#pragma unroll
        for (int i = 0; i < L; ++i) bq[i] = __builtin_nontemporal_load(&vwg[vlo + uint32_t(i) * vrs]);
    } else {
        // last segments: rows at or past nq read row nq - 1 (in bounds) and are zeroed
        uint32_t const vl0 = vlo - uint32_t(r0) * vrs; // row q0 of this line
#pragma unroll
        for (int i = 0; i < L; ++i) {
            double const v = __builtin_nontemporal_load(&vwg[vl0 + uint32_t(min(r0 + i, nq - 1)) * vrs]);
            bq[i] = r0 + i < nq ? v : 0.0;
        }
    }

    const float* T;
    if constexpr (TLDS) {
        // a few float4 per thread (RS <= 10, L = 16, K * L <= 2 * NT rows)
        int const n4 = (K * L * RS) / 4;
        const float4* src = reinterpret_cast<const float4*>(rv.rt);
        float4* dst = reinterpret_cast<float4*>(rt32_lds);
        for (int idx = tid; idx < n4; idx += NT) dst[idx] = src[idx];
        T = rt32_lds + r0 * RS;
    } else {
        T = rv.rt + size_t(r0) * RS;
    }
    if (seg == 0) {
        int64_t const db = lay(dlay, 0, lc);
#pragma unroll
        for (int k = 0; k < kNbMax; ++k) {
            if (k < nb) {
                b2s[k][ln] = active ? rhs_src(bv, vals, vl, vb, dl, dr, dlay, db, k < rv.top ? k : nq + k) : 0.0;
                x2s[k][ln] = 0.0;
            }
        }
    }
    if (tid < LB) lexp[tid] = kNoExp;
    __syncthreads();
    {
        // binary exponent of max |b| over the line
        double m = 0.0;
#pragma unroll
        for (int i = 0; i < L; ++i) m = fmax(m, fabs(bq[i]));
        if (seg == 0)
#pragma unroll
            for (int k = 0; k < kNbMax; ++k)
                if (k < nb) m = fmax(m, fabs(b2s[k][ln]));
        int e = 0;
        (void)frexp(m, &e);
        atomicMax(&lexp[ln], m > 0.0 ? e : kNoExp);
    }
    __syncthreads();
    int const e0 = lexp[ln] == kNoExp ? 0 : lexp[ln];

    // With one refinement step the solution after the first solve is
    // d0 / sc0: an FP32 value times a power of two, held as such (16 VGPRs
    // instead of 32); the FP64 additions are the same, so are the results.
    constexpr bool XF = STEPS == 1;
    double x[XF ? 1 : L];
    float x0[XF ? L : 1];
    double isc0 = 0.0, isc_last = 0.0;
    if constexpr (!XF) {
#pragma unroll
        for (int i = 0; i < L; ++i) x[i] = 0.0;
    }
    auto xval = [&](int i) -> double {
        if constexpr (XF) return double(x0[i]) * isc0;
        else return x[i];
    };
    float d[L];

    // a compile-time step count lets the steps unroll, which keeps the
    // register allocation of the loop body from spilling
#pragma unroll
    for (int it = 0; it <= STEPS; ++it) {
        // ---------------- scaled residual sc (b - A x) ----------------
        int const ex = min(max(it * rv.shift - e0, -1000), 1000);
        double const sc = ldexp(1.0, ex), isc = ldexp(1.0, -ex);
        isc_last = isc;
        if (it == 0) {
#pragma unroll
            for (int i = 0; i < L; ++i) d[i] = float(bq[i] * sc);
            if (seg == 0)
#pragma unroll
                for (int k = 0; k < kNbMax; ++k)
                    if (k < nb) r2s[k][ln] = b2s[k][ln];
        } else {
#pragma unroll
            for (int k = 0; k < KQ; ++k) {
                xch[k][seg][ln] = xval(k);
                xch[KQ + k][seg][ln] = xval(L - KQ + k);
            }
            for (int z = 0; z < rv.bl_nnz; ++z) {
                int const c = rv.bl_c[z];
                if (c >= r0 && c < r0 + L) {
                    double v = 0.0;
#pragma unroll
                    for (int i = 0; i < L; ++i)
                        if (r0 + i == c) v = xval(i);
                    lamx[z][ln] = v;
                }
            }
            __syncthreads();
            int const g0 = rv.gam_seg[seg], g1 = rv.gam_seg[seg + 1];
#pragma unroll
            for (int i = 0; i < L; ++i) {
                const double* Q = rv.rq + size_t(r0 + i) * W;
                double t = bq[i];
#pragma unroll
                for (int j = 0; j < W; ++j) {
                    int const c = i - KQ + j;
                    double v;
                    if (c < 0) v = seg > 0 ? xch[KQ + KQ + c][seg - 1][ln] : 0.0;
                    else if (c >= L) v = seg + 1 < K ? xch[c - L][seg + 1][ln] : 0.0;
                    else v = xval(c);
                    t = fma(-Q[j], v, t);
                }
                for (int z = g0; z < g1; ++z)
                    if (rv.gam_r[z] == r0 + i) t = fma(-rv.gam_v[z], x2s[rv.gam_c[z]][ln], t);
                d[i] = float(t * sc);
                chunk_fence<2>(i);
            }
            if (seg == 0) {
                double t2[kNbMax];
#pragma unroll
                for (int k = 0; k < kNbMax; ++k) t2[k] = k < nb ? b2s[k][ln] : 0.0;
                for (int z = 0; z < rv.bl_nnz; ++z) {
                    int const rr = rv.bl_r[z];
                    double const v = rv.bl_v[z] * lamx[z][ln];
#pragma unroll
                    for (int k = 0; k < kNbMax; ++k)
                        if (k == rr) t2[k] -= v;
                }
#pragma unroll
                for (int k = 0; k < kNbMax; ++k)
#pragma unroll
                    for (int c = 0; c < kNbMax; ++c)
                        if (k < nb && c < nb) t2[k] -= rv.delta[k * nb + c] * x2s[c][ln];
#pragma unroll
                for (int k = 0; k < kNbMax; ++k)
                    if (k < nb) r2s[k][ln] = t2[k];
            }
            __syncthreads(); // halo reads of xch end before the sweeps reuse it
        }

        // ---------------- FP32 forward sweep ----------------
        {
            float s[KF];
#pragma unroll
            for (int k = 0; k < KF; ++k) s[k] = 0.0f;
#pragma unroll
            for (int i = 0; i < L; ++i) {
                const float* R = T + i * RS;
                float t = d[i];
#pragma unroll
                for (int k = KF; k >= 1; --k) t = fmaf(-s[k - 1], R[oLf + k - 1], t);
                if constexpr (FDIV) t = t * R[orDf];
                d[i] = t;
#pragma unroll
                for (int k = KF - 1; k > 0; --k) s[k] = s[k - 1];
                s[0] = t;
                chunk_fence<4>(i);
            }
#pragma unroll
            for (int k = 0; k < KF; ++k) xch[k][seg][ln] = s[k];
            __syncthreads();
            float in[KF];
#pragma unroll
            for (int k = 0; k < KF; ++k) in[k] = 0.0f;
            for (int w = 1; w <= rv.wf; ++w) {
                int const src = seg - w;
                if (src >= 0) {
                    const float* mm = rv.Mf + (size_t(seg) * rv.wf + (w - 1)) * (KF * KF);
#pragma unroll
                    for (int a = 0; a < KF; ++a)
#pragma unroll
                        for (int b = 0; b < KF; ++b) in[a] = fmaf(mm[a * KF + b], float(xch[b][src][ln]), in[a]);
                }
            }
            if (seg > 0) {
#pragma unroll
                for (int i = 0; i < L; ++i) {
                    const float* R = T + i * RS;
                    float t = 0.0f;
#pragma unroll
                    for (int k = KF; k >= 1; --k) t = fmaf(-in[k - 1], R[oLf + k - 1], t);
                    if constexpr (FDIV) t = t * R[orDf];
                    d[i] += t;
#pragma unroll
                    for (int k = KF - 1; k > 0; --k) in[k] = in[k - 1];
                    in[0] = t;
                    chunk_fence<4>(i);
                }
            }
            __syncthreads();
        }

        // ---------------- FP32 backward sweep ----------------
        {
            float s[KB];
#pragma unroll
            for (int k = 0; k < KB; ++k) s[k] = 0.0f;
#pragma unroll
            for (int i = L - 1; i >= 0; --i) {
                const float* R = T + i * RS;
                float t;
                if constexpr (BDIVF) {
                    t = fmaf(-s[0], R[oUb], d[i] * R[orDb]);
                } else {
                    t = d[i];
#pragma unroll
                    for (int k = KB; k >= 1; --k) t = fmaf(-s[k - 1], R[oUb + k - 1], t);
                    t = t * R[orDb];
                }
                d[i] = t;
#pragma unroll
                for (int k = KB - 1; k > 0; --k) s[k] = s[k - 1];
                s[0] = t;
                chunk_fence<4>(i);
            }
#pragma unroll
            for (int k = 0; k < KB; ++k) xch[k][seg][ln] = s[k];
            __syncthreads();
            float in[KB];
#pragma unroll
            for (int k = 0; k < KB; ++k) in[k] = 0.0f;
            for (int w = 1; w <= rv.wb; ++w) {
                int const src = seg + w;
                if (src < K) {
                    const float* mm = rv.Mb + (size_t(seg) * rv.wb + (w - 1)) * (KB * KB);
#pragma unroll
                    for (int a = 0; a < KB; ++a)
#pragma unroll
                        for (int b = 0; b < KB; ++b) in[a] = fmaf(mm[a * KB + b], float(xch[b][src][ln]), in[a]);
                }
            }
            if (seg + 1 < K) {
#pragma unroll
                for (int i = L - 1; i >= 0; --i) {
                    const float* R = T + i * RS;
                    float t;
                    if constexpr (BDIVF) {
                        t = -in[0] * R[oUb];
                    } else {
                        t = 0.0f;
#pragma unroll
                        for (int k = KB; k >= 1; --k) t = fmaf(-in[k - 1], R[oUb + k - 1], t);
                        t = t * R[orDb];
                    }
                    d[i] += t;
#pragma unroll
                    for (int k = KB - 1; k > 0; --k) in[k] = in[k - 1];
                    in[0] = t;
                    chunk_fence<4>(i);
                }
            }
        }

        // ---------------- border and Schur steps (FP64) ----------------
        for (int z = 0; z < rv.bl_nnz; ++z) {
            int const c = rv.bl_c[z];
            if (c >= r0 && c < r0 + L) {
                float v = 0.0f;
#pragma unroll
                for (int i = 0; i < L; ++i)
                    if (r0 + i == c) v = d[i];
                lamx[z][ln] = double(v);
            }
        }
        __syncthreads();
        if (seg == 0) {
            double b2[kNbMax];
#pragma unroll
            for (int k = 0; k < kNbMax; ++k) b2[k] = k < nb ? r2s[k][ln] * sc : 0.0;
            for (int z = 0; z < rv.bl_nnz; ++z) {
                int const rr = rv.bl_r[z];
                double const v = rv.bl_v[z] * lamx[z][ln];
#pragma unroll
                for (int k = 0; k < kNbMax; ++k)
                    if (k == rr) b2[k] -= v;
            }
            getrs_small(nb, rv.s_lu, rv.s_ipiv, b2);
#pragma unroll
            for (int k = 0; k < kNbMax; ++k) {
                if (k < nb) {
                    e2s[k][ln] = b2[k];
                    x2s[k][ln] = x2s[k][ln] + b2[k] * isc;
                }
            }
        }
        __syncthreads();
        int const gblk = rv.gseg[seg];
        if (gblk >= 0) { // Q^-1 gamma is nonzero only in the segments near the block ends
            float e2f[kNbMax];
#pragma unroll
            for (int k = 0; k < kNbMax; ++k) e2f[k] = k < nb ? float(e2s[k][ln]) : 0.0f;
            const float* Gb = rv.Gblk + size_t(gblk) * L * nb;
#pragma unroll
            for (int i = 0; i < L; ++i) {
                float t = d[i];
#pragma unroll
                for (int k = 0; k < kNbMax; ++k)
                    if (k < nb) t = fmaf(-Gb[i * nb + k], e2f[k], t);
                d[i] = t;
            }
        }
        if constexpr (XF) {
            if (it == 0) {
#pragma unroll
                for (int i = 0; i < L; ++i) x0[i] = d[i];
                isc0 = isc;
            }
        } else {
#pragma unroll
            for (int i = 0; i < L; ++i) x[i] = x[i] + double(d[i]) * isc;
        }
    }
    // the solution: x + d / sc of the last step
    auto xout = [&](int i) -> double {
        if constexpr (XF) return double(x0[i]) * isc0 + double(d[i]) * isc_last;
        else return x[i];
    };

    // ---------------- store ----------------
    if (!active) return;
    int const off = bv.offset;
    int const nbas = bv.nbasis;
    uint32_t const c0 = clo + uint32_t(rv.top + r0 + off) * crs;
    if (r0 + L <= nq) {
#pragma unroll
        for (int i = 0; i < L; ++i) __builtin_nontemporal_store(xout(i), &cwg[c0 + uint32_t(i) * crs]);
    } else {
#pragma unroll
        for (int i = 0; i < L; ++i)
            if (r0 + i < nq) __builtin_nontemporal_store(xout(i), &cwg[c0 + uint32_t(i) * crs]);
    }
    if (bv.periodic && (r0 < bv.degree || r0 + L > nbas - off)) {
#pragma unroll
        for (int i = 0; i < L; ++i) {
            int const prow = rv.top + r0 + i;
            if (prow < nq) {
                int const j0 = prow + off - nbas;
                if (j0 >= 0) cwg[clo + uint32_t(j0) * crs] = xout(i);
                int const j2 = prow + off + nbas;
                if (j2 < nbas + bv.degree) cwg[clo + uint32_t(j2) * crs] = xout(i);
            }
        }
    }
    if (seg == 0) {
#pragma unroll
        for (int k = 0; k < kNbMax; ++k) {
            if (k < nb) {
                double const v = x2s[k][ln];
                int const prow = k < rv.top ? k : nq + k;
                coef[cb + int64_t(prow + off) * cl.rs] = v;
                if (bv.periodic) {
                    int const j0 = prow + off - nbas;
                    if (j0 >= 0) coef[cb + int64_t(j0) * cl.rs] = v;
                    int const j2 = prow + off + nbas;
                    if (j2 < nbas + bv.degree) coef[cb + int64_t(j2) * cl.rs] = v;
                }
            }
        }
    }
}

template <int K, int L, int KF, int KB, bool FDIV, bool BDIVF, bool TLDS, int STEPS>
void launch_one(BuilderPlan const& p, size_t lds, double* coef, Layout cl, const double* vals, Layout vl,
                const double* dl, const double* dr, Layout dlay, int64_t nbatch, hipStream_t stream)
{
    constexpr int NT = 1024, LB = NT / K;
    auto kern = build_refine_kernel<LB, K, L, KF, KB, FDIV, BDIVF, TLDS, STEPS>;
    if (lds > 0)
        hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                      int(lds)),
                  "hipFuncSetAttribute");
    int64_t const blocks = (nbatch + LB - 1) / LB;
    hipLaunchKernelGGL(kern, dim3(unsigned(blocks)), dim3(NT), lds, stream, p.rview, p.view, vals, vl, dl, dr, dlay,
                       coef, cl, nbatch);
    hip_check(hipGetLastError(), "build_refine_kernel launch");
}

template <int K, int L, int KF, int KB, bool FDIV, bool BDIVF>
void launch_cfg(BuilderPlan const& p, double* coef, Layout cl, const double* vals, Layout vl, const double* dl,
                const double* dr, Layout dlay, int64_t nbatch, hipStream_t stream)
{
    if (p.refine.rs != KF + KB + 1 + (FDIV ? 1 : 0))
        throw Error(DDCX_ERR_INVALID, "row table record length does not match the band shape");
    constexpr int NT = 1024, LB = NT / K;
    constexpr int KE = KF > KB ? KF : KB;
    constexpr int KX = 2 * KF > KE ? 2 * KF : KE;
    constexpr size_t static_lds =
        size_t(KX * K * LB + kLamMax * LB + 4 * kNbMax * LB) * sizeof(double) + LB * sizeof(int);
    size_t const table = size_t(K) * L * p.refine.rs * sizeof(float);
    bool const tlds = static_lds + table <= size_t(160 * 1024) && p.refine.rs <= 10;
    auto go = [&]<bool TL, int S>() {
        launch_one<K, L, KF, KB, FDIV, BDIVF, TL, S>(p, TL ? table : 0, coef, cl, vals, vl, dl, dr, dlay, nbatch,
                                                     stream);
    };
    bool const one = p.refine.steps == 1;
    if (tlds) {
        if (one) go.template operator()<true, 1>();
        else go.template operator()<true, 2>();
    } else {
        if (one) go.template operator()<false, 1>();
        else go.template operator()<false, 2>();
    }
}

template <int K, int L>
bool launch_K(BuilderPlan const& p, double* coef, Layout cl, const double* vals, Layout vl, const double* dl,
              const double* dr, Layout dlay, int64_t nbatch, hipStream_t stream)
{
    TiledHost const& t = p.refine.t;
    auto go = [&]<int KF, int KB, bool FD, bool BD>() {
        launch_cfg<K, L, KF, KB, FD, BD>(p, coef, cl, vals, vl, dl, dr, dlay, nbatch, stream);
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

} // namespace refine
} // namespace ddcx
