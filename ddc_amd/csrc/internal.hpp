// Internal types of libddcx: host-side spline plans (built and factored once,
// in FP64, on the host) and the plain-old-data views of them that kernels take
// by value.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ddcx.h"
#include "basis.hpp"

namespace ddcx {

// Error carried back across the C ABI as a status code plus message.
struct Error : std::runtime_error {
    ddcx_status status;
    Error(ddcx_status s, std::string const& m) : std::runtime_error(m), status(s) {}
};

void hip_check(hipError_t e, const char* what);

// RAII device buffer (hipMalloc'd once at plan creation).
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(DeviceBuffer const&) = delete;
    DeviceBuffer& operator=(DeviceBuffer const&) = delete;
    ~DeviceBuffer();
    void upload(const void* host, size_t nbytes);
    template <class T>
    void upload(std::vector<T> const& v)
    {
        upload(v.data(), v.size() * sizeof(T));
    }
    template <class T>
    T* as() const
    {
        return static_cast<T*>(ptr);
    }
};

// ---------------------------------------------------------------------------
// Host linear problems (mirror of SplinesLinearProblem*, splines_linear_problem*.cpp)
// ---------------------------------------------------------------------------

enum class QType : int { Dense = 0, Band = 1, PDSBand = 2, PDSTridiag = 3 };

// A square banded or dense matrix, factored on the host with the LAPACK
// algorithms the reference calls (dgbtrf/dpbtrf/dpttrf/dgetrf).
struct HostQ {
    QType type = QType::Dense;
    int n = 0;
    int kl = 0, ku = 0; // Band: kl, ku; PDSBand: kd in kl
    std::vector<double> q; // storage as in the reference (row-major)
    std::vector<int> ipiv; // 0-based
    void init(QType t, int n_, int kl_, int ku_);
    void set(int i, int j, double v);
    double get(int i, int j) const;
    int factor(); // returns LAPACK info
    void solve(double* b, long ld, bool transpose) const; // one right-hand side, stride ld
};

// The spline matrix: a band (or dense) main block Q bordered by nb = n - nq
// rows/columns (SplinesLinearProblem2x2Blocks / 3x3Blocks), or a plain dense
// matrix when the maker selects SplinesLinearProblemDense.
struct HostProblem {
    int n = 0;
    int top = 0;    // 3x3 top size (0 for 2x2)
    bool blocked = false;
    HostQ q;        // main block (whole matrix when !blocked)
    HostQ schur;    // dense nb x nb Schur complement delta - lambda Q^-1 gamma
    std::vector<double> gamma; // nq x nb: Q^-1 gamma after setup
    std::vector<double> lambda; // nb x nq
    // COO copies (|a| >= 1e-14, row-major scan) used by the solve
    std::vector<int> tr_r, tr_c, bl_r, bl_c;
    std::vector<double> tr_v, bl_v;
    // the assembled matrix before setup(): main block, gamma (nq x nb) and
    // delta (nb x nb), for residuals b - A x
    HostQ q_orig;
    std::vector<double> gamma_orig, delta_orig;

    int nq() const { return blocked ? q.n : n; }
    int nb() const { return blocked ? n - q.n : 0; }
    int rhs_rows() const { return n + top; }
    void set(int i, int j, double v);
    double get(int i, int j) const;
    void setup();
    // Solve in the 3x3 index space on a column with rhs_rows() entries.
    void solve(double* b, long ld, bool transpose) const;
};

// ---------------------------------------------------------------------------
// Plans
// ---------------------------------------------------------------------------

// Device copies are made at creation, on the device current then, when a HIP
// device is present; without one the plans serve host-only queries (knots,
// points, quadrature weights, validation errors) and every device entry point
// reports DDCX_ERR_HIP.
struct BSplinesPlan {
    BSplinesView view{}; // host view (knots -> host vector)
    std::vector<double> knots_host;
    DeviceBuffer knots_dev;
    // cell-search brackets of the device view (non-uniform bases with fewer
    // than 65536 cells; see BSplinesView::cells); the host view searches the
    // whole range
    std::vector<uint32_t> cells_host;
    double inv_bucket = 0.0;
    DeviceBuffer cells_dev;
    std::vector<double> integrals; // integrals() over the full spline domain (size() entries)
    DeviceBuffer integrals_dev;
    bool on_device = false; // uploaded at creation when a HIP device is present
    void upload();
    void ensure_device() const; // throws unless the device copies exist
    int nbasis() const { return view.ncells + (view.periodic ? 0 : view.degree); }
    int size() const { return view.ncells + view.degree; }
    BSplinesView device_view() const
    {
        BSplinesView v = view;
        v.knots = view.uniform ? nullptr : knots_dev.as<double>();
        v.cells = cells_host.empty() ? nullptr : cells_dev.as<uint32_t>();
        v.nbucket = int(cells_host.size());
        v.inv_bucket = inv_bucket;
        return v;
    }
};

// Device view of a factored spline matrix, passed by value to the build kernels.
struct SolveView {
    int qtype;       // QType
    int nq, kl, ku;  // main block size and bandwidths
    int nb, top;     // border size, 3x3 top size
    int tr_nnz, bl_nnz;
    const double* q;    // main block factor storage
    const int* ipiv;    // main block pivots
    const double* s_lu; // Schur LU (nb x nb row-major)
    const int* s_ipiv;
    const int* tr_r;
    const int* tr_c;
    const double* tr_v;
    const int* bl_r;
    const int* bl_c;
    const double* bl_v;
};

// SplinesLinearProblemMaker (splines_linear_problem_maker.hpp:56-146) on host
// problems (host_setup.cpp)
void make_band_q(HostQ& q, int n, int kl, int ku, bool pds);
void make_block_problem(HostProblem& pr, int n, int kl, int ku, bool pds, int bottom, int top);
void make_periodic_band_problem(HostProblem& pr, int n, int kl, int ku, bool pds);
void make_dense_block_problem(HostProblem& pr, int n, int bottom, int top);

// Device copy of a factored HostProblem.  upload() fills every field of the
// SolveView: sizes, and pointers into these buffers.
struct SolveBuffers {
    DeviceBuffer q, ipiv, slu, sipiv, trr, trc, trv, blr, blc, blv;
    void upload(HostProblem const& pr, SolveView& sv);
};

// SplinesLinearProblem<ExecSpace> (splines_linear_problem.hpp:26-127) built by
// SplinesLinearProblemMaker (splines_linear_problem_maker.hpp:25-170), for the
// solver entry points of the C ABI: set_element before setup, factor on the
// host in FP64 at setup, then device solves of many right-hand sides.
struct ProblemPlan {
    HostProblem prob;
    bool factored = false;
    bool on_device = false;
    SolveBuffers buf;
    SolveView sv{};
};

struct BuildView {
    SolveView sv;
    int degree, periodic, nbasis, size, offset, npts;
    int bcl, bcu, nbe_l, nbe_r, odd;
    double dxpow[kMaxDegree + 1]; // dx^k, k = 0..degree
};

// Plan of the register-resident tiled build kernel (build_tiled_kernel.hpp).  The
// main block is swept as K segments of L rows per line.  Every sweep is in
// row-oriented form (the same row operations as the netlib column-oriented
// sweeps the reference runs), so the state crossing a segment boundary is the
// last kf solution values (forward) or the next kb ones (backward).  Each
// segment is swept once with a zero incoming state; the incoming state is then
// the short sum
//   in(s) = sum_{w=1..W} P_{s,w} out_loc(s -/+ w)
// of batch-independent transfer products, W chosen so that the first
// neglected product is below a tolerance (2^-60 for the FP64 kernel), and the
// true rows are the local ones plus the homogeneous recurrence of the sweep
// started from that state.  Divisions by pivots are multiplications by
// reciprocals computed here.
struct TiledHost {
    bool ok = false;
    std::string why; // reason when !ok
    int K = 0, L = 0, kf = 0, kb = 0, wf = 0, wb = 0;
    int fwd_div = 0;      // forward rows are scaled by rDf (PDSBand)
    int bwd_divfirst = 0; // backward row form y*rD - x*e (PDSTridiag) instead of (y - sum)*rD
    // per row (padded to K*L rows): coefficients and reciprocal pivots
    std::vector<double> Lf, rDf; // Lf: kf per row
    std::vector<double> Ub, rDb; // Ub: kb per row
    std::vector<double> Mf, Mb;  // K x W x k x k transfer products
    // The sweep tables interleaved per row, rs doubles per row:
    // [Lf(kf) | Ub(kb) | rDb | rDf]
    int rs = 0;
    std::vector<double> rt;
    // Q^-1 gamma (|a| < 1e-14 zeroed) is nonzero only near the ends of the
    // main block: gseg[s] is the index of segment s's L x nb block in Gblk,
    // or -1 when the segment has no nonzero entry.
    std::vector<int> gseg;
    std::vector<double> Gblk;
};

struct TiledView {
    int K, L, kf, kb, wf, wb, nq, nb, top, q0, bl_nnz, rs;
    int ngb; // number of L x nb blocks in Gblk
    const double* rt;
    const double* Mf;
    const double* Mb;
    const int* gseg;
    const double* Gblk;
    const int* bl_r;
    const int* bl_c;
    const double* bl_v;
    const double* s_lu;
    const int* s_ipiv;
};

// Plan of the mixed-precision build (build_refine_kernel.hpp): the tiled sweeps run
// in FP32 on a K x L tiling with FP32 tables (carries to 2^-30), the border
// and Schur steps in FP64, and each line is refined with FP64 residuals
// r = b - A x of the assembled matrix,
//   x_0 = S(b),  x_{k+1} = x_k + S(r_k),  k < steps,
// where S is the approximate solve.  The right-hand side of each S is scaled
// per line by a power of two, 2^(k*shift - e) in step k with e the exponent of
// max |b| over the line and shift = floor(-log2 rho), so the FP32 sweeps see
// values near magnitude 1 whatever the data's range.  steps is the smallest
// count (<= 2) for which the contraction rho measured on the host, with a
// margin of 2, gives (2 rho)^(steps+1) <= 1e-12; a matrix that needs more
// is built by the FP64 kernels.
struct RefineHost {
    bool ok = false;
    std::string why;
    TiledHost t;                // tables in FP64 before rounding
    int rs = 0;                 // floats per row of rt32 (the TiledHost row layout)
    std::vector<float> rt32;    // interleaved row table rounded to FP32
    std::vector<float> Mf32, Mb32, Gblk32;
    int kq = 0;                 // half bandwidth of the main block (kl == ku)
    std::vector<double> rq;     // (K*L) x (2kq+1): main block rows A(u, u-kq+j), zero padded
    std::vector<int> gam_r, gam_c;  // gamma couplings, COO sorted by row
    std::vector<double> gam_v;
    std::vector<int> gam_seg;     // K+1 offsets into the gamma COO, by segment
    std::vector<double> delta;  // nb x nb, row-major
    int steps = 0, shift = 0;
    double rho = 0;
};

struct RefineView {
    int K, L, kf, kb, wf, wb, nq, nb, top, q0, bl_nnz, rs, kq, steps, shift;
    const float* rt;
    const float* Mf;
    const float* Mb;
    const int* gseg;
    const float* Gblk;
    const double* rq;
    const int* gam_r;
    const int* gam_c;
    const double* gam_v;
    const int* gam_seg;
    const double* delta;
    const int* bl_r;
    const int* bl_c;
    const double* bl_v;
    const double* s_lu;
    const int* s_ipiv;
};

struct BuilderPlan {
    std::shared_ptr<BSplinesPlan> bs;
    ddcx_closure bcl, bcu;
    ddcx_solver solver;
    ddcx_precision precision;
    int refine_request = 0; // requested refinement steps, 0 = automatic
    int nbe_l = 0, nbe_r = 0, nbv_l = 0, nbv_r = 0, odd = 0;
    int offset = 0;
    double dx = 0;
    std::vector<double> points;
    HostProblem prob;
    SolveBuffers d_solve;
    DeviceBuffer d_points; // the interpolation points, for the in-kernel feet of ddcx_advect
    BuildView view{}; // device pointers filled by upload()
    TiledHost tiled;
    DeviceBuffer t_Mf, t_Mb, t_rt, t_gseg, t_Gblk;
    TiledView tview{};
    // tables of the fused build + eval kernel (fused_build_eval.hip)
    TiledHost tiled_f;
    DeviceBuffer tf_Mf, tf_Mb, tf_rt, tf_gseg, tf_Gblk;
    TiledView tview_f{};
    // tables of the transposed solve (build_vjp_tiled.hip)
    TiledHost tiled_t;
    DeviceBuffer tt_Mf, tt_Mb, tt_rt, tt_gseg, tt_Gblk;
    TiledView tview_t{};
    RefineHost refine;
    DeviceBuffer r_rt, r_Mf, r_Mb, r_gseg_blk, r_Gblk, r_rq, r_gr, r_gc, r_gv, r_gamseg, r_delta;
    RefineView rview{};
    bool on_device = false;
    void upload();
    void ensure_device() const; // throws unless the device copies exist
};

// segments holding Q^-1 gamma blocks that the tiled adjoint can reduce over
constexpr int kMaxGBlocksVjp = 8;

void make_tiled(BuilderPlan& p);
void make_refine(BuilderPlan& p);

// Batched layout: element (row r, line l) at r*rs + (l % inner)*is + (l / inner)*os.
struct Layout {
    int64_t rs, inner, is, os;
};

inline Layout to_layout(ddcx_layout const& l)
{
    return Layout{l.row_stride, l.inner_count > 0 ? l.inner_count : 1, l.inner_stride, l.outer_stride};
}

// true when the HIP runtime reports at least one device (checked once)
bool device_present();

// host setup entry points (host_setup.cpp)
std::shared_ptr<BSplinesPlan> make_bsplines(ddcx_bsplines_desc const& d);
std::unique_ptr<BuilderPlan> make_builder(std::shared_ptr<BSplinesPlan> bs, ddcx_builder_desc const& d);

// kernel launchers (build_generic.hip / build_tiled_kernel.hpp / eval_kernels.hip)
void launch_build_generic(BuilderPlan const& p, double* coef, Layout cl, const double* vals, Layout vl,
                          const double* dl, const double* dr, Layout dlay, int64_t nbatch, hipStream_t stream);
bool launch_build_tiled(BuilderPlan const& p, double* coef, Layout cl, const double* vals, Layout vl,
                        const double* dl, const double* dr, Layout dlay, int64_t nbatch, hipStream_t stream);
bool launch_build_refine(BuilderPlan const& p, double* coef, Layout cl, const double* vals, Layout vl,
                         const double* dl, const double* dr, Layout dlay, int64_t nbatch, hipStream_t stream);
bool launch_build_vjp_tiled(BuilderPlan const& p, const double* gcoef, Layout cl, double* gvals, Layout vl,
                            double* gdl, double* gdr, Layout dlay, int64_t nbatch, hipStream_t stream);
// Fused build + value-form evaluation; coef may be null (coefficients kept in
// LDS only).  Returns false when the fused kernel does not cover the case.
// With feet_shift (one value per line), coords is not read: point i of line l
// is the interpolation point x_i minus feet_shift[l], and npts is the
// builder's number of interpolation points.
bool launch_build_eval_fused(BuilderPlan const& p, BSplinesPlan const& bs, Rules1D rules, double* coef,
                             Layout cl, const double* vals, Layout vl, const double* dl, const double* dr,
                             Layout dlay, double* out, const double* coords, Layout pl, int64_t npts, int64_t nbatch,
                             hipStream_t stream, const double* feet_shift = nullptr);
void launch_build(BuilderPlan const& p, double* coef, Layout cl, const double* vals, Layout vl, const double* dl,
                  const double* dr, Layout dlay, int64_t nbatch, hipStream_t stream);
// SplinesLinearProblem::solve(b, transpose) on the device: b holds nrhs
// right-hand sides, row i of right-hand side j at b[i * ldb + j]
// (build_generic.hip).
void launch_problem_solve(SolveView const& sv, double* b, int64_t ldb, int64_t nrhs, bool transpose,
                          hipStream_t stream);
void launch_build_vjp(BuilderPlan const& p, const double* gcoef, Layout cl, double* gvals, Layout vl, double* gdl,
                      double* gdr, Layout dlay, int64_t nbatch, hipStream_t stream);
// coordinates at coords + lay(xl, point, line), values at out + lay(pl, point, line)
bool launch_eval_tiled(BSplinesPlan const& bs, Rules1D rules, int order, double* out, const double* coords,
                       Layout xl, Layout pl, int64_t npts, const double* coef, Layout cl, int64_t nbatch,
                       hipStream_t stream);
bool launch_eval_vjp_tiled(BSplinesPlan const& bs, Rules1D rules, const double* coords, const double* w,
                           double* gx, Layout pl, int64_t npts, const double* coef, double* gcoef, Layout cl,
                           int64_t nbatch, hipStream_t stream);
void launch_eval(BSplinesPlan const& bs, Rules1D rules, int order, double* out, const double* coords,
                 Layout pl, int64_t npts, const double* coef, Layout cl, int64_t nbatch, hipStream_t stream);
// the coordinates in a layout of their own (xl), e.g. one vector shared by every line
void launch_eval_xl(BSplinesPlan const& bs, Rules1D rules, int order, double* out, const double* coords,
                    Layout xl, Layout pl, int64_t npts, const double* coef, Layout cl, int64_t nbatch,
                    hipStream_t stream);
void launch_eval_vjp(BSplinesPlan const& bs, Rules1D rules, const double* coords, const double* w,
                     double* gx, Layout pl, int64_t npts, const double* coef, double* gcoef, Layout cl,
                     int64_t nbatch, hipStream_t stream);
void launch_integrate(BSplinesPlan const& bs, double* out, int64_t ostride, const double* coef, Layout cl,
                      int64_t nbatch, hipStream_t stream);
// N-dimensional tensor-product evaluation and integration (eval_nd.hip)
constexpr int kMaxNd = 4;
void launch_evalnd(BSplinesPlan const* const* bs, int nd, const int* orders, const int* lower, const int* upper,
                   double* out, const double* coords, int64_t npts, const double* coef, int64_t cbs,
                   const int64_t* st, int64_t nbatch, hipStream_t stream);
// binned 3D evaluation (eval_nd.hip): workspace bytes (0: not supported for
// this grid, use launch_evalnd), and the launch
int64_t evalnd_binned_workspace(BSplinesPlan const* const* bs, int nd, int64_t npts, int64_t nbatch);
void launch_evalnd_binned(BSplinesPlan const* const* bs, int nd, const int* orders, const int* lower,
                          const int* upper, double* out, const double* coords, int64_t npts, const double* coef,
                          int64_t cbs, const int64_t* st, int64_t nbatch, void* work, int64_t work_bytes,
                          hipStream_t stream);

// compute units of the current device (cached per device)
int cu_count();

// workgroups per batch entry of the integration (1: no partial sums)
int integratend_split(BSplinesPlan const* const* bs, int nd, int64_t nbatch);
// part: nbatch * S doubles of partial sums when S > 1
void launch_integratend(BSplinesPlan const* const* bs, int nd, double* out, int64_t ostride, const double* coef,
                        int64_t cbs, const int64_t* st, int64_t nbatch, double* part, int S, hipStream_t stream);
void launch_eval2d(BSplinesPlan const& b1, BSplinesPlan const& b2, Rules2D const& rules, int d1, int d2, double* out,
                   const double* coords, int64_t npts, const double* coef, int64_t coef_batch_stride, int64_t st1,
                   int64_t st2, int64_t nbatch, hipStream_t stream);
// binned 2D adjoint (eval2d_binned.hip): workspace bytes, 0 if unsupported
int64_t eval2d_vjp_workspace(BSplinesPlan const& bp1, BSplinesPlan const& bp2, int64_t npts, int64_t nbatch);
// packed 2D evaluation (eval2d_packed.hip): workspace bytes per batch entry
int64_t eval2d_packed_record_bytes(BSplinesPlan const& bp1, BSplinesPlan const& bp2);
void launch_eval2d_packed(BSplinesPlan const& bp1, BSplinesPlan const& bp2, Rules2D const& rules, int o1, int o2,
                          double* out, const double* coords, int64_t npts, const double* coef, int64_t cbs,
                          int64_t st1, int64_t st2, int64_t nbatch, void* work, int64_t work_bytes,
                          hipStream_t stream);
// device address of the current GPU's copy of the static workspace
// (DDCX_STATIC_SCRATCH_BYTES), resolved once per device
void* static_scratch();
void launch_eval2d_vjp_binned(BSplinesPlan const& bp1, BSplinesPlan const& bp2, Rules2D const& rules,
                              const double* coords, const double* w, double* gx, int64_t npts, const double* coef,
                              double* gcoef, int64_t cbs, int64_t st1, int64_t st2, int64_t nbatch, void* work,
                              int64_t work_bytes, hipStream_t stream);
void launch_eval2d_vjp(BSplinesPlan const& b1, BSplinesPlan const& b2, Rules2D const& rules, const double* coords,
                       const double* w, double* gx, int64_t npts, const double* coef, double* gcoef,
                       int64_t coef_batch_stride, int64_t st1, int64_t st2, int64_t nbatch, hipStream_t stream);

} // namespace ddcx
