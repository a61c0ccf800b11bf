// Generic batched spline build: one thread per batch line, sweeping the
// factored matrix in the reference's operation order (dpttrs / dpbtrs /
// dgbtrs / dgetrs restated, then the 2x2-block corrections of
// splines_linear_problem_2x2_blocks.cpp:275-298).  It handles every matrix
// shape the maker can produce and every layout; the register-resident tiled
// kernel (build_tiled_kernel.hpp) takes the hot configurations.
//
// Data flow per line (spline_builder.hpp:896-1101 without the separate copy,
// transpose and duplication passes): the right-hand side is read straight
// from vals / derivs, the forward sweep writes y into the coefficient array at
// its final rows, the backward sweep turns y into x in place, the border rows
// are solved in registers, and the periodic duplicate rows are written last.
#include "internal.hpp"

namespace ddcx {

namespace {

__device__ __forceinline__ int64_t lay(Layout const& L, int64_t r, int64_t l)
{
    return r * L.rs + (l % L.inner) * L.is + (l / L.inner) * L.os;
}

struct LineIO {
    BuildView const& bv;
    const double* vals;
    const double* dl;
    const double* dr;
    Layout vl, dlay;
    double* coef; // coef + column offset of this line
    int64_t crs;
    int64_t vbase, dbase;

    // physical right-hand-side row i (0 <= i < nbasis): the value the
    // reference places at spline(i + offset) before the solve (:949-1033)
    __device__ double src(int i) const
    {
        if (bv.periodic) return vals[vbase + int64_t(i) * vl.rs];
        if (i < bv.nbe_l) {
            if (bv.bcl == DDCX_HERMITE) return dl[dbase + int64_t(i) * dlay.rs] * bv.dxpow[i + bv.odd];
            return 0.0;
        }
        int const nv_end = bv.nbasis - bv.nbe_r;
        if (i < nv_end) return vals[vbase + int64_t(i - bv.nbe_l) * vl.rs];
        if (bv.bcu == DDCX_HERMITE) return dr[dbase + int64_t(i - nv_end) * dlay.rs] * bv.dxpow[i - nv_end + bv.odd];
        return 0.0;
    }
    __device__ double& c(int phys_row) const { return coef[int64_t(phys_row + bv.offset) * crs]; }
};

// ---- forward + backward sweeps over the main block Q (rows top .. top+nq-1) ----

template <class IO>
__device__ void sweep_tridiag(SolveView const& sv, IO const& io)
{
    int const n = sv.nq, t = sv.top;
    const double* d = sv.q;
    const double* e = sv.q + n;
    if (n == 1) {
        io.c(t) = io.src(t) * (1.0 / d[0]);
        return;
    }
    double y = io.src(t);
    io.c(t) = y;
    for (int i = 1; i < n; ++i) {
        y = io.src(t + i) - y * e[i - 1];
        io.c(t + i) = y;
    }
    double x = y / d[n - 1];
    io.c(t + n - 1) = x;
    for (int i = n - 2; i >= 0; --i) {
        x = io.c(t + i) / d[i] - x * e[i];
        io.c(t + i) = x;
    }
}

template <int K, class IO>
__device__ void sweep_pdsband(SolveView const& sv, IO const& io)
{
    int const n = sv.nq, t = sv.top;
    const double* q = sv.q;
    double w[K + 1];
#pragma unroll
    for (int k = 0; k <= K; ++k) w[k] = k < n ? io.src(t + k) : 0.0;
    for (int j = 0; j < n; ++j) { // dtbsv('L','N','N')
        double const xj = w[0] / q[j];
        io.c(t + j) = xj;
#pragma unroll
        for (int i = 1; i <= K; ++i)
            if (j + i < n) w[i] = w[i] - xj * q[size_t(i) * n + j];
#pragma unroll
        for (int k = 0; k < K; ++k) w[k] = w[k + 1];
        w[K] = (j + K + 1 < n) ? io.src(t + j + K + 1) : 0.0;
    }
    double xw[K > 0 ? K : 1]; // xw[i-1] = x[j+i]
#pragma unroll
    for (int k = 0; k < (K > 0 ? K : 1); ++k) xw[k] = 0.0;
    for (int j = n - 1; j >= 0; --j) { // dtbsv('L','T','N')
        double temp = io.c(t + j);
#pragma unroll
        for (int i = K; i >= 1; --i)
            if (j + i < n) temp = temp - q[size_t(i) * n + j] * xw[i - 1];
        double const xj = temp / q[j];
        io.c(t + j) = xj;
#pragma unroll
        for (int k = (K > 0 ? K : 1) - 1; k > 0; --k) xw[k] = xw[k - 1];
        xw[0] = xj;
    }
}

template <int KL, int KU, class IO>
__device__ void sweep_band(SolveView const& sv, IO const& io)
{
    constexpr int KD = KL + KU;
    int const n = sv.nq, t = sv.top;
    const double* q = sv.q;
    auto AB = [&](int r, int cidx) { return q[size_t(r) * n + cidx]; };
    // forward: unit lower factor with row interchanges (dgbtrs, trans = 'N')
    double w[KL + 1];
#pragma unroll
    for (int k = 0; k <= KL; ++k) w[k] = k < n ? io.src(t + k) : 0.0;
    for (int j = 0; j < n - 1; ++j) {
        int const lm = min(KL, n - 1 - j);
        int const p = sv.ipiv[j] - j;
#pragma unroll
        for (int k = 1; k <= KL; ++k)
            if (k == p) {
                double const tmp = w[0];
                w[0] = w[k];
                w[k] = tmp;
            }
        double const temp = -1.0 * w[0];
#pragma unroll
        for (int i = 0; i < KL; ++i)
            if (i < lm) w[1 + i] = w[1 + i] + AB(KD + 1 + i, j) * temp;
        io.c(t + j) = w[0];
#pragma unroll
        for (int k = 0; k < KL; ++k) w[k] = w[k + 1];
        w[KL] = (j + KL + 1 < n) ? io.src(t + j + KL + 1) : 0.0;
    }
    io.c(t + n - 1) = w[0];
    // backward: upper band factor, bandwidth kl+ku (dtbsv('U','N','N'))
    double u[KD + 1]; // u[KD - i] = row j - i
#pragma unroll
    for (int i = 0; i <= KD; ++i) u[KD - i] = (n - 1 - i >= 0) ? io.c(t + n - 1 - i) : 0.0;
    for (int j = n - 1; j >= 0; --j) {
        double const xj = u[KD] / AB(KD, j);
        io.c(t + j) = xj;
#pragma unroll
        for (int i = 1; i <= KD; ++i)
            if (j - i >= 0) u[KD - i] = u[KD - i] - xj * AB(KD - i, j);
#pragma unroll
        for (int k = KD; k > 0; --k) u[k] = u[k - 1];
        u[0] = (j - KD - 1 >= 0) ? io.c(t + j - KD - 1) : 0.0;
    }
}

constexpr int kMaxDenseQ = 32;

__device__ void getrs_local(int n, const double* a, const int* ipiv, double* b)
{
    for (int j = 0; j < n; ++j)
        if (ipiv[j] != j) {
            double const tmp = b[ipiv[j]];
            b[ipiv[j]] = b[j];
            b[j] = tmp;
        }
    for (int j = 0; j < n; ++j)
        for (int i = j + 1; i < n; ++i) b[i] = b[i] - b[j] * a[i * n + j];
    for (int j = n - 1; j >= 0; --j) {
        b[j] = b[j] / a[j * n + j];
        for (int i = 0; i < j; ++i) b[i] = b[i] - b[j] * a[i * n + j];
    }
}

template <class IO>
__device__ void sweep_dense(SolveView const& sv, IO const& io)
{
    int const n = sv.nq, t = sv.top;
    double b[kMaxDenseQ];
    for (int i = 0; i < n; ++i) b[i] = io.src(t + i);
    getrs_local(n, sv.q, sv.ipiv, b);
    for (int i = 0; i < n; ++i) io.c(t + i) = b[i];
}

constexpr int kMaxBorder = 16;

// 2x2-block corrections (2x2_blocks.cpp:290-291) on the border rows.
template <class IO>
__device__ void solve_border(SolveView const& sv, IO const& io)
{
    int const nq = sv.nq, nb = sv.nb, t = sv.top;
    if (nb > 0) {
        double b2[kMaxBorder];
        for (int k = 0; k < nb; ++k) b2[k] = io.src(k < t ? k : nq + k);
        for (int z = 0; z < sv.bl_nnz; ++z) b2[sv.bl_r[z]] -= sv.bl_v[z] * io.c(t + sv.bl_c[z]);
        getrs_local(nb, sv.s_lu, sv.s_ipiv, b2);
        int z = 0;
        while (z < sv.tr_nnz) {
            int const r = sv.tr_r[z];
            double x = io.c(t + r);
            for (; z < sv.tr_nnz && sv.tr_r[z] == r; ++z) x -= sv.tr_v[z] * b2[sv.tr_c[z]];
            io.c(t + r) = x;
        }
        for (int k = 0; k < nb; ++k) io.c(k < t ? k : nq + k) = b2[k];
    }
}

// The border solve, then the periodic duplication (spline_builder.hpp:1074-1100).
__device__ void border_and_dup(BuildView const& bv, LineIO const& io)
{
    solve_border(bv.sv, io);
    if (bv.periodic) {
        int const off = bv.offset, nbas = bv.nbasis;
        double* base = io.coef;
        int64_t const rs = io.crs;
        if (off != 0) {
            for (int i = 0; i < off; ++i) base[int64_t(i) * rs] = base[int64_t(nbas + i) * rs];
            for (int i = off; i < bv.degree; ++i) base[int64_t(nbas + i) * rs] = base[int64_t(i) * rs];
        }
        for (int i = 0; i < bv.degree; ++i) base[int64_t(nbas + i) * rs] = base[int64_t(i) * rs];
    }
}

template <class Sweep>
__global__ void __launch_bounds__(256) build_generic_kernel(BuildView bv, const double* __restrict__ vals, Layout vl,
                                                            const double* __restrict__ dl,
                                                            const double* __restrict__ dr, Layout dlay,
                                                            double* __restrict__ coef, Layout cl, int64_t nbatch)
{
    int64_t const l = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (l >= nbatch) return;
    LineIO io{bv, vals, dl, dr, vl, dlay, coef + lay(cl, 0, l), cl.rs, lay(vl, 0, l), lay(dlay, 0, l)};
    Sweep::run(bv.sv, io);
    border_and_dup(bv, io);
}

struct TridiagSweep {
    template <class IO>
    static __device__ void run(SolveView const& sv, IO const& io) { sweep_tridiag(sv, io); }
};
template <int K>
struct PDSBandSweep {
    template <class IO>
    static __device__ void run(SolveView const& sv, IO const& io) { sweep_pdsband<K>(sv, io); }
};
template <int KL, int KU>
struct BandSweep {
    template <class IO>
    static __device__ void run(SolveView const& sv, IO const& io) { sweep_band<KL, KU>(sv, io); }
};
struct DenseSweep {
    template <class IO>
    static __device__ void run(SolveView const& sv, IO const& io) { sweep_dense(sv, io); }
};

template <class Sweep>
void launch_with(BuildView const& bv, double* coef, Layout cl, const double* vals, Layout vl, const double* dl,
                 const double* dr, Layout dlay, int64_t nbatch, hipStream_t stream)
{
    int const threads = 256;
    int64_t const blocks = (nbatch + threads - 1) / threads;
    hipLaunchKernelGGL(build_generic_kernel<Sweep>, dim3(unsigned(blocks)), dim3(threads), 0, stream, bv, vals, vl,
                       dl, dr, dlay, coef, cl, nbatch);
    hip_check(hipGetLastError(), "build_generic_kernel launch");
}

} // namespace

void launch_build_generic(BuilderPlan const& p, double* coef, Layout cl, const double* vals, Layout vl,
                          const double* dl, const double* dr, Layout dlay, int64_t nbatch, hipStream_t stream)
{
    SolveView const& sv = p.view.sv;
    if (sv.nb > kMaxBorder) throw Error(DDCX_ERR_UNSUPPORTED, "border block larger than 16 rows");
    auto const qt = QType(sv.qtype);
    auto go = [&]<class S>() { launch_with<S>(p.view, coef, cl, vals, vl, dl, dr, dlay, nbatch, stream); };
    switch (qt) {
    case QType::PDSTridiag: go.template operator()<TridiagSweep>(); return;
    case QType::Dense:
        if (sv.nq > kMaxDenseQ) throw Error(DDCX_ERR_UNSUPPORTED, "dense main block larger than 32 rows");
        go.template operator()<DenseSweep>();
        return;
    case QType::PDSBand:
        switch (sv.kl) {
        case 0: go.template operator()<PDSBandSweep<0>>(); return;
        case 1: go.template operator()<PDSBandSweep<1>>(); return;
        case 2: go.template operator()<PDSBandSweep<2>>(); return;
        case 3: go.template operator()<PDSBandSweep<3>>(); return;
        default: break;
        }
        break;
    case QType::Band:
        if (sv.kl == sv.ku) {
            switch (sv.kl) {
            case 0: go.template operator()<BandSweep<0, 0>>(); return;
            case 1: go.template operator()<BandSweep<1, 1>>(); return;
            case 2: go.template operator()<BandSweep<2, 2>>(); return;
            case 3: go.template operator()<BandSweep<3, 3>>(); return;
            case 4: go.template operator()<BandSweep<4, 4>>(); return;
            case 5: go.template operator()<BandSweep<5, 5>>(); return;
            case 6: go.template operator()<BandSweep<6, 6>>(); return;
            default: break;
            }
        }
        break;
    }
    throw Error(DDCX_ERR_UNSUPPORTED, "no build kernel for this band shape");
}

} // namespace ddcx

namespace ddcx {

void launch_build(BuilderPlan const& p, double* coef, Layout cl, const double* vals, Layout vl, const double* dl,
                  const double* dr, Layout dlay, int64_t nbatch, hipStream_t stream)
{
    p.ensure_device();
    if (p.precision == DDCX_FP32_REFINE && launch_build_refine(p, coef, cl, vals, vl, dl, dr, dlay, nbatch, stream))
        return;
    if (launch_build_tiled(p, coef, cl, vals, vl, dl, dr, dlay, nbatch, stream)) return;
    launch_build_generic(p, coef, cl, vals, vl, dl, dr, dlay, nbatch, stream);
}

namespace {

// ---------------------------------------------------------------------------
// Adjoint of the build: the transposed solve of the reference
// (SplinesLinearProblem::solve(b, true); 2x2_blocks.cpp:292-297 order), one
// thread per line, working in place in the gradient-of-values array.
// ---------------------------------------------------------------------------

struct VjpIO {
    BuildView const& bv;
    const double* gcoef;
    Layout cl;
    int64_t cb;
    double* gvals;
    int64_t vrs, vb;
    int q0;
    // folded gradient of the physical unknown i: the sum over every
    // coefficient row j with (j - offset) mod nbasis == i, in increasing j
    __device__ double fold(int i) const
    {
        if (!bv.periodic) return gcoef[cb + int64_t(i) * cl.rs];
        int const j1 = i + bv.offset, n = bv.nbasis, N = bv.size;
        double acc = 0.0;
        if (j1 - n >= 0) acc += gcoef[cb + int64_t(j1 - n) * cl.rs];
        acc += gcoef[cb + int64_t(j1) * cl.rs];
        if (j1 + n < N) acc += gcoef[cb + int64_t(j1 + n) * cl.rs];
        return acc;
    }
    __device__ double& w(int r) const { return gvals[vb + int64_t(q0 + r) * vrs]; }
};

__device__ void getrs_trans_local(int n, const double* a, const int* ipiv, double* b)
{
    for (int j = 0; j < n; ++j) {
        double temp = b[j];
        for (int i = 0; i < j; ++i) temp = temp - a[i * n + j] * b[i];
        b[j] = temp / a[j * n + j];
    }
    for (int j = n - 1; j >= 0; --j) {
        double temp = b[j];
        for (int i = j + 1; i < n; ++i) temp = temp - a[i * n + j] * b[i];
        b[j] = temp;
    }
    for (int j = n - 1; j >= 0; --j)
        if (ipiv[j] != j) {
            double const tmp = b[ipiv[j]];
            b[ipiv[j]] = b[j];
            b[j] = tmp;
        }
}

// Q^T solve in place on the main-block rows (PDS types are symmetric: the
// reference ignores `transpose` for them).  io.w(r) is main-block row r.
template <class IO>
__device__ void q_solve_trans(SolveView const& sv, IO const& io)
{
    int const n = sv.nq;
    switch (QType(sv.qtype)) {
    case QType::PDSTridiag: {
        const double* d = sv.q;
        const double* e = sv.q + n;
        if (n == 1) {
            io.w(0) = io.w(0) * (1.0 / d[0]);
            return;
        }
        for (int i = 1; i < n; ++i) io.w(i) = io.w(i) - io.w(i - 1) * e[i - 1];
        io.w(n - 1) = io.w(n - 1) / d[n - 1];
        for (int i = n - 2; i >= 0; --i) io.w(i) = io.w(i) / d[i] - io.w(i + 1) * e[i];
        return;
    }
    case QType::PDSBand: {
        int const k = sv.kl;
        const double* q = sv.q;
        for (int j = 0; j < n; ++j) {
            double const xj = io.w(j) / q[j];
            io.w(j) = xj;
            for (int i = j + 1; i <= min(n - 1, j + k); ++i) io.w(i) = io.w(i) - xj * q[size_t(i - j) * n + j];
        }
        for (int j = n - 1; j >= 0; --j) {
            double temp = io.w(j);
            for (int i = min(n - 1, j + k); i >= j + 1; --i) temp = temp - q[size_t(i - j) * n + j] * io.w(i);
            io.w(j) = temp / q[j];
        }
        return;
    }
    case QType::Band: {
        int const kl = sv.kl, kd = sv.kl + sv.ku;
        const double* q = sv.q;
        auto AB = [&](int r, int c) { return q[size_t(r) * n + c]; };
        for (int j = 0; j < n; ++j) { // dtbsv('U','T','N')
            double temp = io.w(j);
            for (int i = max(0, j - kd); i <= j - 1; ++i) temp = temp - AB(kd + i - j, j) * io.w(i);
            io.w(j) = temp / AB(kd, j);
        }
        for (int j = n - 2; j >= 0; --j) { // L^T with the interchanges in reverse
            int const lm = min(kl, n - 1 - j);
            double temp = 0.0;
            for (int i = 0; i < lm; ++i) temp = temp + AB(kd + 1 + i, j) * io.w(j + 1 + i);
            io.w(j) = io.w(j) + -1.0 * temp;
            int const l = sv.ipiv[j];
            if (l != j) {
                double const t = io.w(l);
                io.w(l) = io.w(j);
                io.w(j) = t;
            }
        }
        return;
    }
    case QType::Dense: {
        double b[kMaxDenseQ];
        for (int i = 0; i < n; ++i) b[i] = io.w(i);
        getrs_trans_local(n, sv.q, sv.ipiv, b);
        for (int i = 0; i < n; ++i) io.w(i) = b[i];
        return;
    }
    }
}

__global__ void __launch_bounds__(256) build_vjp_generic_kernel(BuildView bv, const double* __restrict__ gcoef,
                                                                Layout cl, double* __restrict__ gvals, Layout vl,
                                                                double* __restrict__ gdl, double* __restrict__ gdr,
                                                                Layout dlay, int64_t nbatch)
{
    int64_t const l = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (l >= nbatch) return;
    SolveView const& sv = bv.sv;
    int const nq = sv.nq, nb = sv.nb, t = sv.top;
    int const q0 = bv.periodic ? 0 : t - bv.nbe_l;
    VjpIO io{bv, gcoef, cl, lay(cl, 0, l), gvals, vl.rs, lay(vl, 0, l), q0};
    for (int r = 0; r < nq; ++r) io.w(r) = io.fold(t + r);
    double b2[kMaxBorder];
    for (int k = 0; k < nb; ++k) b2[k] = io.fold(k < t ? k : nq + k);
    if (nb > 0) {
        for (int z = 0; z < sv.tr_nnz; ++z) b2[sv.tr_c[z]] -= sv.tr_v[z] * io.w(sv.tr_r[z]);
        getrs_trans_local(nb, sv.s_lu, sv.s_ipiv, b2);
        for (int z = 0; z < sv.bl_nnz; ++z) io.w(sv.bl_c[z]) -= sv.bl_v[z] * b2[sv.bl_r[z]];
    }
    q_solve_trans(sv, io);
    // border rows back to the value / derivative gradients they came from
    int64_t const db = lay(dlay, 0, l);
    int const nv_end = bv.nbasis - bv.nbe_r;
    for (int k = 0; k < nb; ++k) {
        int const i = k < t ? k : nq + k;
        if (bv.periodic) {
            io.gvals[io.vb + int64_t(i) * vl.rs] = b2[k];
        } else if (i < bv.nbe_l) {
            if (bv.bcl == DDCX_HERMITE) gdl[db + int64_t(i) * dlay.rs] = b2[k] * bv.dxpow[i + bv.odd];
        } else if (i < nv_end) {
            io.gvals[io.vb + int64_t(i - bv.nbe_l) * vl.rs] = b2[k];
        } else if (bv.bcu == DDCX_HERMITE) {
            gdr[db + int64_t(i - nv_end) * dlay.rs] = b2[k] * bv.dxpow[i - nv_end + bv.odd];
        }
    }
}

} // namespace

void launch_build_vjp(BuilderPlan const& p, const double* gcoef, Layout cl, double* gvals, Layout vl, double* gdl,
                      double* gdr, Layout dlay, int64_t nbatch, hipStream_t stream)
{
    p.ensure_device();
    if (launch_build_vjp_tiled(p, gcoef, cl, gvals, vl, gdl, gdr, dlay, nbatch, stream)) return;
    SolveView const& sv = p.view.sv;
    if (sv.nb > kMaxBorder) throw Error(DDCX_ERR_UNSUPPORTED, "border block larger than 16 rows");
    if (QType(sv.qtype) == QType::Dense && sv.nq > kMaxDenseQ)
        throw Error(DDCX_ERR_UNSUPPORTED, "dense main block larger than 32 rows");
    int const threads = 256;
    int64_t const blocks = (nbatch + threads - 1) / threads;
    hipLaunchKernelGGL(build_vjp_generic_kernel, dim3(unsigned(blocks)), dim3(threads), 0, stream, p.view, gcoef, cl,
                       gvals, vl, gdl, gdr, dlay, nbatch);
    hip_check(hipGetLastError(), "build_vjp_generic_kernel launch");
}

} // namespace ddcx

// ---------------------------------------------------------------------------
// SplinesLinearProblem::solve(b, transpose) on its own
// (splines_linear_problem.hpp:77): one thread per right-hand side, in place,
// through the same sweeps as the builds.  Row i of right-hand side j is
// b[i * ldb + j], the reference's LayoutRight MultiRHS view (rows, nrhs).
// ---------------------------------------------------------------------------

namespace ddcx {

namespace {

struct ColumnIO {
    double* b;
    int64_t ldb;
    int top;
    __device__ double src(int i) const { return b[int64_t(i) * ldb]; }
    __device__ double& c(int i) const { return b[int64_t(i) * ldb]; }
    __device__ double& w(int r) const { return b[int64_t(top + r) * ldb]; } // main-block row r
};

// dgbtrs('N') with run-time bandwidths, in place: the band shapes the spline
// builder never produces (kl != ku, or kl > 6).
struct BandSweepRT {
    static __device__ void run(SolveView const& sv, ColumnIO const& io)
    {
        int const n = sv.nq, t = sv.top, kl = sv.kl, kd = sv.kl + sv.ku;
        const double* q = sv.q;
        auto AB = [&](int r, int cidx) { return q[size_t(r) * n + cidx]; };
        for (int j = 0; j < n - 1; ++j) {
            int const lm = min(kl, n - 1 - j);
            int const p = sv.ipiv[j];
            if (p != j) {
                double const tmp = io.c(t + p);
                io.c(t + p) = io.c(t + j);
                io.c(t + j) = tmp;
            }
            double const temp = -1.0 * io.c(t + j);
            for (int i = 0; i < lm; ++i) io.c(t + j + 1 + i) = io.c(t + j + 1 + i) + AB(kd + 1 + i, j) * temp;
        }
        for (int j = n - 1; j >= 0; --j) { // dtbsv('U','N','N')
            double const xj = io.c(t + j) / AB(kd, j);
            io.c(t + j) = xj;
            for (int i = 1; i <= kd && j - i >= 0; ++i) io.c(t + j - i) = io.c(t + j - i) - xj * AB(kd - i, j);
        }
    }
};

template <class Sweep>
__global__ void __launch_bounds__(256) problem_solve_kernel(SolveView sv, double* __restrict__ b, int64_t ldb,
                                                            int64_t nrhs)
{
    int64_t const j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= nrhs) return;
    ColumnIO const io{b + j, ldb, sv.top};
    Sweep::run(sv, io);
    solve_border(sv, io);
}

// the transposed order of 2x2_blocks.cpp:292-297 (and HostProblem::solve)
__global__ void __launch_bounds__(256) problem_solve_trans_kernel(SolveView sv, double* __restrict__ b, int64_t ldb,
                                                                  int64_t nrhs)
{
    int64_t const j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= nrhs) return;
    ColumnIO const io{b + j, ldb, sv.top};
    int const nq = sv.nq, nb = sv.nb, t = sv.top;
    double b2[kMaxBorder];
    for (int k = 0; k < nb; ++k) b2[k] = io.src(k < t ? k : nq + k);
    if (nb > 0) {
        for (int z = 0; z < sv.tr_nnz; ++z) b2[sv.tr_c[z]] -= sv.tr_v[z] * io.w(sv.tr_r[z]);
        getrs_trans_local(nb, sv.s_lu, sv.s_ipiv, b2);
        for (int z = 0; z < sv.bl_nnz; ++z) io.w(sv.bl_c[z]) -= sv.bl_v[z] * b2[sv.bl_r[z]];
    }
    q_solve_trans(sv, io);
    for (int k = 0; k < nb; ++k) io.c(k < t ? k : nq + k) = b2[k];
}

} // namespace

void launch_problem_solve(SolveView const& sv, double* b, int64_t ldb, int64_t nrhs, bool transpose,
                          hipStream_t stream)
{
    if (sv.nb > kMaxBorder) throw Error(DDCX_ERR_UNSUPPORTED, "border block larger than 16 rows");
    auto const qt = QType(sv.qtype);
    if (qt == QType::Dense && sv.nq > kMaxDenseQ) throw Error(DDCX_ERR_UNSUPPORTED, "dense block larger than 32 rows");
    if (nrhs == 0) return;
    int const threads = 256;
    dim3 const grid(unsigned((nrhs + threads - 1) / threads));
    if (transpose) {
        hipLaunchKernelGGL(problem_solve_trans_kernel, grid, dim3(threads), 0, stream, sv, b, ldb, nrhs);
        hip_check(hipGetLastError(), "problem_solve_trans_kernel launch");
        return;
    }
    auto go = [&]<class S>() {
        hipLaunchKernelGGL(problem_solve_kernel<S>, grid, dim3(threads), 0, stream, sv, b, ldb, nrhs);
        hip_check(hipGetLastError(), "problem_solve_kernel launch");
    };
    switch (qt) {
    case QType::PDSTridiag: go.template operator()<TridiagSweep>(); return;
    case QType::Dense: go.template operator()<DenseSweep>(); return;
    case QType::PDSBand:
        switch (sv.kl) {
        case 1: go.template operator()<PDSBandSweep<1>>(); return;
        case 2: go.template operator()<PDSBandSweep<2>>(); return;
        case 3: go.template operator()<PDSBandSweep<3>>(); return;
        case 4: go.template operator()<PDSBandSweep<4>>(); return;
        case 5: go.template operator()<PDSBandSweep<5>>(); return;
        case 6: go.template operator()<PDSBandSweep<6>>(); return;
        default: throw Error(DDCX_ERR_UNSUPPORTED, "positive-definite band solves support 1 to 6 off-diagonals");
        }
    case QType::Band:
        if (sv.kl == sv.ku) {
            switch (sv.kl) {
            case 1: go.template operator()<BandSweep<1, 1>>(); return;
            case 2: go.template operator()<BandSweep<2, 2>>(); return;
            case 3: go.template operator()<BandSweep<3, 3>>(); return;
            case 4: go.template operator()<BandSweep<4, 4>>(); return;
            case 5: go.template operator()<BandSweep<5, 5>>(); return;
            case 6: go.template operator()<BandSweep<6, 6>>(); return;
            default: break;
            }
        }
        go.template operator()<BandSweepRT>();
        return;
    }
    throw Error(DDCX_ERR_UNSUPPORTED, "no solve kernel for this matrix shape");
}

} // namespace ddcx
