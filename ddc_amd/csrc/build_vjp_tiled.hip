// Register-resident tiled adjoint of the batched build: the transposed solve
// of SplinesLinearProblem2x2Blocks::solve(b, true)
// (splines_linear_problem_2x2_blocks.cpp:275-298: the direct order reversed)
// after the fold of the periodic duplicate rows (the transpose of
// spline_builder.hpp:1074-1100), on the tiling of build_tiled_kernel.hpp:
//   1. y = folded gradient of the main-block rows (L per thread), b2 = the
//      folded gradient of the border rows;
//   2. b2 -= (Q^-1 gamma)^T y: partial sums of the segments that hold Q^-1
//      gamma blocks, combined in segment order by one thread per line;
//   3. b2 = S^-T b2 (Schur complement LU, transposed);
//   4. y -= lambda^T b2;
//   5. y = Q^-T y with the tiled sweeps (transposed tables);
//   6. the rows back to the value and Hermite-derivative gradients.
#include "tiled_common.hpp"

namespace ddcx {

namespace {

using namespace tiled;

// dgetrs('T') on an nb x nb LU (nb <= kNbMax) held in registers
__device__ __forceinline__ void getrs_trans_small(int nb, const double* a, const int* ipiv, double* b)
{
#pragma unroll
    for (int j = 0; j < kNbMax; ++j) {
        if (j < nb) {
            double t = b[j];
#pragma unroll
            for (int i = 0; i < j; ++i) t = t - a[i * nb + j] * b[i];
            b[j] = t / a[j * nb + j];
        }
    }
#pragma unroll
    for (int j = kNbMax - 1; j >= 0; --j) {
        if (j < nb) {
            double t = b[j];
#pragma unroll
            for (int i = j + 1; i < kNbMax; ++i)
                if (i < nb) t = t - a[i * nb + j] * b[i];
            b[j] = t;
        }
    }
#pragma unroll
    for (int j = kNbMax - 1; j >= 0; --j) {
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
}

template <int LB, int K, int L, int KF, int KB, bool FDIV, bool BDIVF, bool TLDS, bool IC>
__global__ void __launch_bounds__(LB* K) build_vjp_tiled_kernel(TiledView tv, BuildView bv,
                                                               const double* __restrict__ gcoef, Layout cl,
                                                               double* __restrict__ gvals, Layout vl,
                                                               double* __restrict__ gdl, double* __restrict__ gdr,
                                                               Layout dlay, int64_t nbatch)
{
    constexpr int KE = KF > KB ? KF : KB;
    constexpr int NT = LB * K;
    constexpr int SR = stage_rows(K * L), SP = stage_pitch(K * L); // staging chunk rows, pitch (IC)
    __shared__ double ex[KE][K][LB];
    __shared__ double part[kMaxGBlocksVjp][kNbMax][LB];
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
    int64_t const cb = lay(cl, 0, lc);
    int64_t const vb = lay(vl, 0, lc);
    int const r0 = seg * L;
    int const nq = tv.nq, nb = tv.nb, top = tv.top;
    int const off = bv.offset, nbas = bv.nbasis, nsz = bv.size;
    constexpr int RS = KF + KB + 1 + (FDIV ? 1 : 0); // row record length, fixed by the band shape
    const double* __restrict__ cwg = gcoef + lay(cl, 0, line0);
    double* __restrict__ vwg = gvals + lay(vl, 0, line0);
    uint32_t const clo = uint32_t(cb - lay(cl, 0, line0));
    uint32_t const vlo = uint32_t(vb - lay(vl, 0, line0)) + uint32_t(tv.q0 + r0) * uint32_t(vl.rs);
    uint32_t const crs = uint32_t(cl.rs), vrs = uint32_t(vl.rs);

    // dynamic LDS: the row table and the padded Q^-1 gamma blocks (TLDS),
    // then the staging area (IC), LB x stage_pitch(K * L) doubles, and the per-line
    // offsets of gcoef and gvals
    double* gpad = rt_lds + K * L * RS;
    double* stg = rt_lds + (TLDS ? K * L * RS + tv.ngb * L * kNbMax : 0);
    uint32_t* loff = reinterpret_cast<uint32_t*>(stg + LB * SP);
    int const nl = int(min<int64_t>(LB, nbatch - line0));
    if constexpr (IC) {
        if (tid < LB) {
            int64_t const l = line0 + min(tid, nl - 1);
            loff[tid] = uint32_t(lay(cl, 0, l) - lay(cl, 0, line0));
            loff[LB + tid] = uint32_t(lay(vl, 0, l) - lay(vl, 0, line0));
        }
        __syncthreads();
    }

    // 1. folded gradient: unknown i collects the coefficient rows j with
    //    (j - offset) mod nbasis == i (a single row except at the periodic ends)
    double y[L];
    {
        uint32_t const c0 = clo + uint32_t(top + r0 + off) * crs;
        if constexpr (IC) {
            // each line's main-block rows as contiguous runs through LDS
            constexpr int PER = LB * SR / NT;
            for (int rbase = 0; rbase < K * L; rbase += SR) {
                double tmp[PER];
#pragma unroll
                for (int k = 0; k < PER; ++k) {
                    int const e = tid + k * NT;
                    int const l = e / SR, u = rbase + e % SR;
                    tmp[k] = (l < nl && u < nq) ? __builtin_nontemporal_load(&cwg[loff[l] + uint32_t(top + u + off)])
                                                : 0.0;
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
#pragma unroll
            for (int i = 0; i < L; ++i) y[i] = __builtin_nontemporal_load(&cwg[c0 + uint32_t(i) * crs]);
        } else {
            uint32_t const cl0 = clo + uint32_t(top + off) * crs;
#pragma unroll
            for (int i = 0; i < L; ++i) {
                double const v = __builtin_nontemporal_load(&cwg[cl0 + uint32_t(min(r0 + i, nq - 1)) * crs]);
                y[i] = r0 + i < nq ? v : 0.0;
            }
        }
        if (bv.periodic && (r0 < bv.degree || r0 + L > nbas - off)) {
#pragma unroll
            for (int i = 0; i < L; ++i) {
                int const j1 = top + r0 + i + off;
                if (r0 + i < nq) {
                    if (j1 - nbas >= 0) y[i] += cwg[clo + uint32_t(j1 - nbas) * crs];
                    if (j1 + nbas < nsz) y[i] += cwg[clo + uint32_t(j1 + nbas) * crs];
                }
            }
        }
    }

    // the border rows' incoming gradient (folded like the main rows), read
    // now and parked in x2s, which segment 0 overwrites with x2 later
    if (nb > 0 && seg == 0) {
#pragma unroll
        for (int k = 0; k < kNbMax; ++k) {
            if (k < nb) {
                int const j1 = (k < top ? k : nq + k) + off;
                double v = active ? gcoef[cb + int64_t(j1) * cl.rs] : 0.0;
                if (active && bv.periodic) {
                    if (j1 - nbas >= 0) v += gcoef[cb + int64_t(j1 - nbas) * cl.rs];
                    if (j1 + nbas < nsz) v += gcoef[cb + int64_t(j1 + nbas) * cl.rs];
                }
                x2s[k][ln] = v;
            }
        }
    }

    // per-row table and border data in LDS when they fit
    int const gblk = nb > 0 ? tv.gseg[seg] : -1;
    const double* T;
    if constexpr (TLDS) {
        int const n2 = (K * L * RS) / 2;
        const double2* src = reinterpret_cast<const double2*>(tv.rt);
        double2* dst = reinterpret_cast<double2*>(rt_lds);
        for (int idx = tid; idx < n2; idx += NT) dst[idx] = src[idx];
        stage_border<NT, true>(tv, bl, gpad, tid);
        __syncthreads();
        T = rt_lds + r0 * RS;
    } else {
        stage_border<NT, false>(tv, bl, gpad, tid); // couplings only
        T = tv.rt + size_t(r0) * RS;
    }

    // 2. partial sums of (Q^-1 gamma)^T y over this segment's rows
    if (gblk >= 0) {
        if constexpr (TLDS) {
            // padded columns (k >= nb) sum to zero and are never read
            const double* Gb = gpad + gblk * L * kNbMax;
#pragma unroll
            for (int k = 0; k < kNbMax; ++k) {
                double acc = 0.0;
#pragma unroll
                for (int i = 0; i < L; ++i) acc = fma(Gb[i * kNbMax + k], y[i], acc);
                part[gblk][k][ln] = acc;
            }
        } else {
            const double* Gb = tv.Gblk + size_t(gblk) * L * nb;
#pragma unroll
            for (int k = 0; k < kNbMax; ++k) {
                if (k < nb) {
                    double acc = 0.0;
#pragma unroll
                    for (int i = 0; i < L; ++i) acc = fma(Gb[i * nb + k], y[i], acc);
                    part[gblk][k][ln] = acc;
                }
            }
        }
    }
    __syncthreads();

    // 3. the border: b2 -= sum of the partials (segment order), b2 = S^-T b2
    if (seg == 0) {
        double b2[kNbMax];
        int const nblk = tv.ngb;
#pragma unroll
        for (int k = 0; k < kNbMax; ++k) {
            if (k < nb) {
                double v = x2s[k][ln];
                for (int g = 0; g < nblk; ++g) v -= part[g][k][ln];
                b2[k] = v;
            } else {
                b2[k] = 0.0;
            }
        }
        if (nb > 0) getrs_trans_small(nb, tv.s_lu, tv.s_ipiv, b2);
#pragma unroll
        for (int k = 0; k < kNbMax; ++k)
            if (k < nb) x2s[k][ln] = b2[k];
        // 6 (border part). rows back to the gradients they came from
        if (active) {
            int64_t const db = lay(dlay, 0, lc);
            int const nv_end = nbas - bv.nbe_r;
#pragma unroll
            for (int k = 0; k < kNbMax; ++k) {
                if (k < nb) {
                    int const i = k < top ? k : nq + k;
                    if (bv.periodic) {
                        gvals[vb + int64_t(i) * vl.rs] = b2[k];
                    } else if (i < bv.nbe_l) {
                        if (bv.bcl == DDCX_HERMITE) gdl[db + int64_t(i) * dlay.rs] = b2[k] * bv.dxpow[i + bv.odd];
                    } else if (i < nv_end) {
                        gvals[vb + int64_t(i - bv.nbe_l) * vl.rs] = b2[k];
                    } else if (bv.bcu == DDCX_HERMITE) {
                        gdr[db + int64_t(i - nv_end) * dlay.rs] = b2[k] * bv.dxpow[i - nv_end + bv.odd];
                    }
                }
            }
        }
    }
    __syncthreads();

    // 4. y -= lambda^T x2
    for (int z = 0; z < tv.bl_nnz; ++z) {
        int const c = bl.c[z];
        if (c >= r0 && c < r0 + L) {
            double const v = bl.v[z] * x2s[bl.r[z]][ln];
#pragma unroll
            for (int i = 0; i < L; ++i)
                if (r0 + i == c) y[i] -= v;
        }
    }

    // 5. y = Q^-T y
    tiled_sweeps<LB, K, L, KF, KB, FDIV, BDIVF, KE>(y, T, RS, tv, ex, seg, ln);

    // 6. main-block rows to the value gradients
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
                                                &vwg[loff[LB + l] + uint32_t(tv.q0 + u)]);
            }
            __syncthreads();
        }
    } else if (active) {
        if (r0 + L <= nq) {
#pragma unroll
            for (int i = 0; i < L; ++i) __builtin_nontemporal_store(y[i], &vwg[vlo + uint32_t(i) * vrs]);
        } else {
#pragma unroll
            for (int i = 0; i < L; ++i)
                if (r0 + i < nq) __builtin_nontemporal_store(y[i], &vwg[vlo + uint32_t(i) * vrs]);
        }
    }
}

template <int K, int L, int KF, int KB, bool FDIV, bool BDIVF, int NT>
void launch_cfg(BuilderPlan const& p, const double* gcoef, Layout cl, double* gvals, Layout vl, double* gdl,
                double* gdr, Layout dlay, int64_t nbatch, hipStream_t stream)
{
    if (p.tiled_t.rs != KF + KB + 1 + (FDIV ? 1 : 0))
        throw Error(DDCX_ERR_INVALID, "row table record length does not match the band shape");
    constexpr int LB = NT / K;
    constexpr int KE = KF > KB ? KF : KB;
    constexpr size_t static_lds =
        size_t(KE * K * LB + kMaxGBlocksVjp * kNbMax * LB + kNbMax * LB) * sizeof(double) + sizeof(BorderLds);
    size_t const table = (size_t(K) * L * p.tiled_t.rs + gamma_lds_doubles(p.tview_t)) * sizeof(double);
    bool const ic = vl.rs == 1 && cl.rs == 1;
    size_t const stage = ic ? size_t(LB) * stage_pitch(K * L) * sizeof(double) + 2 * LB * sizeof(uint32_t) : 0;
    bool const tlds = static_lds + table + stage <= size_t(160 * 1024) && p.tiled_t.rs <= 10;
    size_t const lds = (tlds ? table : 0) + stage;
    int64_t const blocks = (nbatch + LB - 1) / LB;
    auto go = [&]<bool TL, bool I>() {
        auto kern = build_vjp_tiled_kernel<LB, K, L, KF, KB, FDIV, BDIVF, TL, I>;
        if (lds > 0)
            hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)),
                      "hipFuncSetAttribute");
        hipLaunchKernelGGL(kern, dim3(unsigned(blocks)), dim3(NT), lds, stream, p.tview_t, p.view, gcoef, cl, gvals,
                           vl, gdl, gdr, dlay, nbatch);
    };
    if (ic) {
        if (tlds) go.template operator()<true, true>();
        else go.template operator()<false, true>();
    } else {
        if (tlds) go.template operator()<true, false>();
        else go.template operator()<false, false>();
    }
    hip_check(hipGetLastError(), "build_vjp_tiled_kernel launch");
}

// NT threads per workgroup: NT / K lines of K segments of L rows
template <int K, int L, int NT = 1024>
bool launch_K(BuilderPlan const& p, const double* gcoef, Layout cl, double* gvals, Layout vl, double* gdl,
              double* gdr, Layout dlay, int64_t nbatch, hipStream_t stream)
{
    TiledHost const& t = p.tiled_t;
    auto go = [&]<int KF, int KB, bool FD, bool BD>() {
        launch_cfg<K, L, KF, KB, FD, BD, NT>(p, gcoef, cl, gvals, vl, gdl, gdr, dlay, nbatch, stream);
        return true;
    };
    if (t.bwd_divfirst) {
        if (t.kf == 1 && t.kb == 1 && !t.fwd_div) return go.template operator()<1, 1, false, true>();
        return false;
    }
    if (!t.fwd_div) return false;
    // PDS band, and the transposed unpivoted band
    if (t.kf == 1 && t.kb == 1) return go.template operator()<1, 1, true, false>();
    if (t.kf == 2 && t.kb == 2) return go.template operator()<2, 2, true, false>();
    if (t.kf == 3 && t.kb == 3) return go.template operator()<3, 3, true, false>();
    if (t.kf == 4 && t.kb == 4) return go.template operator()<4, 4, true, false>();
    return false;
}

} // namespace

bool launch_build_vjp_tiled(BuilderPlan const& p, const double* gcoef, Layout cl, double* gvals, Layout vl,
                            double* gdl, double* gdr, Layout dlay, int64_t nbatch, hipStream_t stream)
{
    if (!p.tiled_t.ok) return false;
    if (!offsets_fit(vl, p.tiled_t.K * p.tiled_t.L + p.tview_t.q0, nbatch) || !offsets_fit(cl, p.view.size, nbatch))
        return false;
    if (p.tiled_t.K == 32 && p.tiled_t.L == 32)
        return launch_K<32, 32>(p, gcoef, cl, gvals, vl, gdl, gdr, dlay, nbatch, stream);
    if (p.tiled_t.K == 64 && p.tiled_t.L == 32)
        return launch_K<64, 32>(p, gcoef, cl, gvals, vl, gdl, gdr, dlay, nbatch, stream);
    // short lines
    if (p.tiled_t.K == 16 && p.tiled_t.L == 8)
        return launch_K<16, 8, 512>(p, gcoef, cl, gvals, vl, gdl, gdr, dlay, nbatch, stream);
    if (p.tiled_t.K == 16 && p.tiled_t.L == 9)
        return launch_K<16, 9, 512>(p, gcoef, cl, gvals, vl, gdl, gdr, dlay, nbatch, stream);
    if (p.tiled_t.K == 32 && p.tiled_t.L == 8)
        return launch_K<32, 8>(p, gcoef, cl, gvals, vl, gdl, gdr, dlay, nbatch, stream);
    return false;
}

} // namespace ddcx
