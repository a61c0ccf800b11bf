// Fused batched build + evaluation: SplineBuilder::operator()
// (spline_builder.hpp:905) followed by SplineEvaluator::operator()
// (spline_evaluator.hpp:263) on the same lines, in one kernel.  A workgroup
// solves LB lines with the register-resident tiled sweeps (build_tiled_kernel.hpp,
// 16-row segments), writes the coefficients into an LDS tile that reuses the
// row table's space, and evaluates every point of those lines from it
// (eval_tiled.hip's batch-contiguous mapping).  The coefficients go to HBM only
// when the caller asks for them, so a step reads vals and coords and writes
// the values: 24 bytes per point-and-row instead of the 40 of the two calls.
#include "tiled_common.hpp"

namespace ddcx {

namespace {

using namespace tiled;

template <int LB, int K, int L, int KF, int KB, bool FDIV, bool BDIVF, int D>
__global__ void __launch_bounds__(LB* K) fused_build_eval_kernel(TiledView tv, BuildView bv, BSplinesView b,
                                                                Rules1D rules, const double* __restrict__ vals,
                                                                Layout vl, const double* __restrict__ dl,
                                                                const double* __restrict__ dr, Layout dlay,
                                                                double* __restrict__ coef, Layout cl,
                                                                double* __restrict__ out,
                                                                const double* __restrict__ coords, Layout pl,
                                                                int64_t npts, int64_t nbatch,
                                                                const double* __restrict__ fpts,
                                                                const double* __restrict__ fshift)
{
    constexpr int KE = KF > KB ? KF : KB;
    constexpr int NT = LB * K;
    constexpr int ROWS = NT / LB; // point rows per pass (= K)
    constexpr int U = 8;
    __shared__ double ex[KE][K][LB];
    __shared__ double lamx[kLamMax][LB];
    __shared__ double x2s[kNbMax][LB];
    __shared__ BorderLds bl;
    // row table and padded Q^-1 gamma blocks during the build, coefficient
    // tile (size x LB) after
    extern __shared__ double dyn[];

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
    const double* __restrict__ vwg = vals + lay(vl, 0, line0);
    uint32_t const vlo = uint32_t(vb - lay(vl, 0, line0)) + uint32_t(tv.q0 + r0) * uint32_t(vl.rs);
    uint32_t const vrs = uint32_t(vl.rs);
    bool const store = coef != nullptr;

    // ---------------- build ----------------
    double y[L];
    if (r0 + L <= nq) {
#pragma unroll
        for (int i = 0; i < L; ++i) y[i] = __builtin_nontemporal_load(&vwg[vlo + uint32_t(i) * vrs]);
    } else {
        uint32_t const vl0 = vlo - uint32_t(r0) * vrs; // row q0 of this line
#pragma unroll
        for (int i = 0; i < L; ++i) {
            double const v = __builtin_nontemporal_load(&vwg[vl0 + uint32_t(min(r0 + i, nq - 1)) * vrs]);
            y[i] = r0 + i < nq ? v : 0.0;
        }
    }
    // the border rows' right-hand side, loaded with the tile and parked in
    // x2s (segment 0 overwrites it with the border solution later)
    if (tv.nb > 0 && seg == 0) {
        int64_t const db = lay(dlay, 0, lc);
#pragma unroll
        for (int k = 0; k < kNbMax; ++k)
            if (k < tv.nb)
                x2s[k][ln] = active ? rhs_src(bv, vals, vl, vb, dl, dr, dlay, db, k < tv.top ? k : nq + k) : 0.0;
    }
    int const gblk = tv.nb > 0 ? tv.gseg[seg] : -1; // Q^-1 gamma block (-1: none)
    double* const gpad = dyn + K * L * RS;
    {
        int const n2 = (K * L * RS) / 2;
        const double2* src = reinterpret_cast<const double2*>(tv.rt);
        double2* dst = reinterpret_cast<double2*>(dyn);
        for (int idx = tid; idx < n2; idx += NT) dst[idx] = src[idx];
        stage_border<NT, true>(tv, bl, gpad, tid);
        __syncthreads();
    }
    tiled_sweeps<LB, K, L, KF, KB, FDIV, BDIVF, KE>(y, dyn + r0 * RS, RS, tv, ex, seg, ln);

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
            for (int k = 0; k < kNbMax; ++k)
                if (k < nb) x2s[k][ln] = b2[k];
        }
        __syncthreads();
        if (gblk >= 0) {
            double x2[kNbMax];
#pragma unroll
            for (int k = 0; k < kNbMax; ++k) x2[k] = k < nb ? x2s[k][ln] : 0.0;
            apply_gamma<L>(y, gpad + gblk * L * kNbMax, x2);
        }
    }
    __syncthreads(); // the row table is dead from here on

    // ---------------- coefficients: LDS tile (and HBM when asked) ----------------
    double* tile = dyn;
    auto put = [&](int row, double v) {
        tile[row * LB + ln] = v;
        if (store && active) __builtin_nontemporal_store(v, &coef[cb + int64_t(row) * cl.rs]);
    };
#pragma unroll
    for (int i = 0; i < L; ++i) {
        int const prow = tv.top + r0 + i;
        if (r0 + i < nq) {
            put(prow + off, y[i]);
            if (bv.periodic) {
                if (prow + off - nbas >= 0) put(prow + off - nbas, y[i]);
                if (prow + off + nbas < nbas + bv.degree) put(prow + off + nbas, y[i]);
            }
        }
    }
    if (seg == 0) {
#pragma unroll
        for (int k = 0; k < kNbMax; ++k) {
            if (k < nb) {
                double const v = x2s[k][ln];
                int const prow = k < tv.top ? k : nq + k;
                put(prow + off, v);
                if (bv.periodic) {
                    if (prow + off - nbas >= 0) put(prow + off - nbas, v);
                    if (prow + off + nbas < nbas + bv.degree) put(prow + off + nbas, v);
                }
            }
        }
    }
    __syncthreads();

    // ---------------- evaluation of every point of the LB lines ----------------
    if (!active) return;
    int64_t const pb = lay(pl, 0, lc);
    const double* tl = tile + ln;
    int const prow0 = seg; // point row of this thread within a pass
    // the point: read from coords, or the foot fpts[i] - fshift[line] of an
    // advection step (npts interpolation points)
    double const sh = fshift ? fshift[lc] : 0.0;
    auto point = [&](int64_t ii) {
        return fshift ? fpts[ii] - sh : __builtin_nontemporal_load(&coords[pb + ii * pl.rs]);
    };
    double xa[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        int64_t const ii = prow0 + int64_t(u) * ROWS;
        xa[u] = ii < npts ? point(ii) : 0.0;
    }
    for (int64_t i = prow0; i < npts; i += U * ROWS) {
        double xb[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            int64_t const ii = i + int64_t(U + u) * ROWS;
            xb[u] = ii < npts ? point(ii) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            int64_t const ii = i + int64_t(u) * ROWS;
            if (ii < npts) {
                double x = xa[u];
                bool outside;
                double yv = 0.0;
                if (apply_bc(b, rules, x, outside)) {
                    double v[D + 1];
                    int const jmin = eval_basis<D, true>(b, b.knots, x, v);
#pragma unroll
                    for (int k = 0; k < D + 1; ++k) yv += tl[(jmin + k) * LB] * v[k];
                }
                __builtin_nontemporal_store(yv, &out[pb + ii * pl.rs]);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) xa[u] = xb[u];
    }
}

template <int K, int KF, int KB, bool FDIV, bool BDIVF, int D>
void launch_one(BuilderPlan const& p, BSplinesView const& b, Rules1D rules, double* coef, Layout cl,
                const double* vals, Layout vl, const double* dl, const double* dr, Layout dlay, double* out,
                const double* coords, Layout pl, int64_t npts, int64_t nbatch, hipStream_t stream, size_t lds,
                const double* fpts, const double* fshift)
{
    if (p.tiled_f.rs != KF + KB + 1 + (FDIV ? 1 : 0))
        throw Error(DDCX_ERR_INVALID, "row table record length does not match the band shape");
    constexpr int NT = 1024, LB = NT / K, L = 16;
    auto kern = fused_build_eval_kernel<LB, K, L, KF, KB, FDIV, BDIVF, D>;
    hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  int(lds)),
              "hipFuncSetAttribute");
    int64_t const blocks = (nbatch + LB - 1) / LB;
    hipLaunchKernelGGL(kern, dim3(unsigned(blocks)), dim3(NT), lds, stream, p.tview_f, p.view, b, rules, vals,
                       vl, dl, dr, dlay, coef, cl, out, coords, pl, npts, nbatch, fpts, fshift);
    hip_check(hipGetLastError(), "fused_build_eval_kernel launch");
}

template <int K>
bool launch_K(BuilderPlan const& p, BSplinesView const& b, Rules1D rules, double* coef, Layout cl,
              const double* vals, Layout vl, const double* dl, const double* dr, Layout dlay, double* out,
              const double* coords, Layout pl, int64_t npts, int64_t nbatch, hipStream_t stream, size_t lds,
              const double* fpts, const double* fshift)
{
    TiledHost const& t = p.tiled_f;
    int const deg = b.degree;
    bool done = false;
    auto go = [&]<int KF, int KB, bool FD, bool BD, int D>() {
        if (!done && deg == D) {
            launch_one<K, KF, KB, FD, BD, D>(p, b, rules, coef, cl, vals, vl, dl, dr, dlay, out, coords, pl,
                                            npts, nbatch, stream, lds, fpts, fshift);
            done = true;
        }
    };
    if (t.bwd_divfirst) { // PDS tridiagonal: uniform periodic of degree 2, 3
        go.template operator()<1, 1, false, true, 2>();
        go.template operator()<1, 1, false, true, 3>();
    } else if (t.fwd_div) { // PDS band: uniform periodic of degree 4, 5
        if (t.kf == 2) {
            go.template operator()<2, 2, true, false, 4>();
            go.template operator()<2, 2, true, false, 5>();
        }
    } else if (t.kf == 1) { // band, kl = degree / 2 (uniform) or degree - 1 (non-uniform)
        go.template operator()<1, 1, false, false, 2>();
        go.template operator()<1, 1, false, false, 3>();
    } else if (t.kf == 2) {
        go.template operator()<2, 2, false, false, 3>();
        go.template operator()<2, 2, false, false, 4>();
        go.template operator()<2, 2, false, false, 5>();
    } else if (t.kf == 3) {
        go.template operator()<3, 3, false, false, 4>();
    } else if (t.kf == 4) {
        go.template operator()<4, 4, false, false, 5>();
    }
    return done;
}

} // namespace

bool launch_build_eval_fused(BuilderPlan const& p, BSplinesPlan const& bs, Rules1D rules, double* coef,
                             Layout cl, const double* vals, Layout vl, const double* dl, const double* dr,
                             Layout dlay, double* out, const double* coords, Layout pl, int64_t npts, int64_t nbatch,
                             hipStream_t stream, const double* feet_shift)
{
    TiledHost const& t = p.tiled_f;
    if (!t.ok || p.bs.get() != &bs) return false;
    if (feet_shift && npts != int64_t(p.points.size())) return false;
    const double* const fpts = feet_shift ? p.d_points.as<double>() : nullptr;
    if (vl.rs == 1 || pl.rs == 1 || (coef && cl.rs == 1)) return false; // batch-contiguous layouts only
    int const K = t.K, L = t.L, LB = 1024 / K;
    if (L != 16 || (K != 64 && K != 128)) return false;
    if (!offsets_fit(vl, int64_t(K) * L + p.tview_f.q0, nbatch)) return false;
    // dynamic LDS: the row table and Q^-1 gamma blocks during the build, then
    // the coefficient tile
    size_t const table = (size_t(K) * L * t.rs + gamma_lds_doubles(p.tview_f)) * sizeof(double);
    size_t const tile = size_t(bs.size()) * LB * sizeof(double);
    size_t const lds = std::max(table, tile);
    int const KE = std::max(t.kf, t.kb);
    size_t const stat = size_t(KE * K * LB + kLamMax * LB + kNbMax * LB) * sizeof(double) + sizeof(BorderLds);
    if (lds + stat > size_t(160 * 1024)) return false;
    BSplinesView const b = bs.device_view();
    if (K == 64)
        return launch_K<64>(p, b, rules, coef, cl, vals, vl, dl, dr, dlay, out, coords, pl, npts, nbatch,
                            stream, lds, fpts, feet_shift);
    return launch_K<128>(p, b, rules, coef, cl, vals, vl, dl, dr, dlay, out, coords, pl, npts, nbatch, stream,
                         lds, fpts, feet_shift);
}

} // namespace ddcx
