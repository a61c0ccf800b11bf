/*
 * ddcx.h -- C ABI of libddcx, the MI355X (gfx950) implementation of DDC's
 * batched spline hot path: ddc::SplineBuilder / ddc::SplineEvaluator (1D and
 * tensor-product 2D) with their closures, solver tags and extrapolation rules.
 *
 * All citations are to DDC v0.15.1 (src/ddc/kernels/splines/...).  The
 * reference exposes this path as C++ templates, not as a C interface; each
 * entry point below names the reference member it stands in for.
 *
 * Conventions
 *  - Plain pointers and sizes only.  Array arguments of build/eval/integrate
 *    calls are DEVICE pointers (hipMalloc'd or equivalent); the `stream`
 *    argument is a hipStream_t (NULL = default stream).  Calls are
 *    stream-ordered and asynchronous; nothing allocates device memory after
 *    create (plans own all device data).  Plans upload their device data at
 *    create (ddcx_problem: at setup), synchronously, to the device current
 *    then; call them with streams of that device.  Without a HIP device the
 *    handles still answer host queries, and device calls return DDCX_ERR_HIP.
 *  - A batched array is described by a ddcx_layout: element (row r, batch
 *    line l) lives at  r*row_stride + (l % inner_count)*inner_stride
 *    + (l / inner_count)*outer_stride  (units: elements).  This covers every
 *    layout_right dimension order of a ddc::ChunkSpan with one interpolation
 *    dimension and up to two batch dimensions.
 *  - Every function returns a ddcx_status; on error ddcx_last_error() gives a
 *    thread-local message (the facade rethrows std::runtime_error with it, as
 *    the reference throws).
 */
#ifndef DDCX_H
#define DDCX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum ddcx_status {
    DDCX_OK = 0,
    DDCX_ERR_INVALID = 1,    /* bad argument / shape (reference: assert) */
    DDCX_ERR_GRID = 2,       /* check_valid_grid, spline_builder.hpp:1273-1317 */
    DDCX_ERR_SINGULAR = 3,   /* LAPACK info != 0, splines_linear_problem_band.cpp:121-123 */
    DDCX_ERR_HIP = 4,        /* HIP runtime error */
    DDCX_ERR_UNSUPPORTED = 5 /* outside this build's range (e.g. degree > 7) */
} ddcx_status;

/* ddc::SplineBuilderClosure, spline_builder_closures.hpp:18-23 (same order) */
typedef enum ddcx_closure {
    DDCX_PERIODIC = 0,
    DDCX_HERMITE = 1,
    DDCX_HOMOGENEOUS_HERMITE = 2,
    DDCX_GREVILLE = 3
} ddcx_closure;

/* ddc::SplineSolver, spline_builder.hpp:36-39.  GINKGO is accepted and routed
 * to the direct (LAPACK-algorithm) solver. */
typedef enum ddcx_solver { DDCX_SOLVER_GINKGO = 0, DDCX_SOLVER_LAPACK = 1 } ddcx_solver;

/* Extrapolation rules: NullExtrapolationRule (null_extrapolation_rule.hpp:14-26),
 * ConstantExtrapolationRule (constant_extrapolation_rule.hpp:27-73),
 * PeriodicExtrapolationRule (periodic_extrapolation_rule.hpp:12-24). */
typedef enum ddcx_extrap { DDCX_EXTRAP_NULL = 0, DDCX_EXTRAP_CONSTANT = 1, DDCX_EXTRAP_PERIODIC = 2 } ddcx_extrap;

/* Arithmetic of the batched solve: plain FP64 (the drop-in), or FP32 sweeps
 * with FP64 residual refinement.  No reference counterpart. */
typedef enum ddcx_precision { DDCX_FP64 = 0, DDCX_FP32_REFINE = 1 } ddcx_precision;

typedef struct ddcx_layout {
    int64_t row_stride;
    int64_t inner_count;
    int64_t inner_stride;
    int64_t outer_stride;
} ddcx_layout;

/* UniformBSplines<CDim, D>::Impl(rmin, rmax, ncells) (bsplines_uniform.hpp:135-148)
 * or NonUniformBSplines<CDim, D>::Impl(breaks) (bsplines_non_uniform.hpp:414-457).
 * periodic = CDim::PERIODIC.  breaks: host pointer, ncells+1 sorted values. */
typedef struct ddcx_bsplines_desc {
    int degree;
    int uniform;
    int periodic;
    int ncells;
    double rmin;
    double rmax;
    const double* breaks;
} ddcx_bsplines_desc;

typedef struct ddcx_bsplines ddcx_bsplines;
typedef struct ddcx_builder ddcx_builder;
typedef struct ddcx_evaluator ddcx_evaluator;

/* SplineBuilder<Exec, Mem, BSplines, DDimI, BCl, BCu, Solver> constructor,
 * spline_builder.hpp:227-260 (with check_valid_grid :1273-1317,
 * compute_offset :627-658, allocate_matrix :719-774, build_matrix_system
 * :784-886).  points: host pointer to the interpolation points (npoints of
 * them), or NULL for GrevilleInterpolationPoints<BSplines, BCl, BCu>
 * (greville_interpolation_points.hpp:153-269). */
typedef struct ddcx_builder_desc {
    ddcx_closure lower;
    ddcx_closure upper;
    ddcx_solver solver;
    ddcx_precision precision;
    const double* points;
    int npoints;
    /* DDCX_FP32_REFINE: refinement steps after the first FP32 solve; 0 picks
     * the smallest count (1 or 2) whose error bound from the contraction
     * measured at create time is <= 1e-12, 1 or 2 requests that count. */
    int refine_steps;
} ddcx_builder_desc;

typedef struct ddcx_builder_info {
    int npoints;      /* interpolation points (rows of vals) */
    int nbasis;       /* UniformBSplines::nbasis(), bsplines_uniform.hpp:339-342 */
    int size;         /* coefficients per line, BSplines::size() */
    int offset;       /* periodic RHS rotation, spline_builder.hpp:627-658 */
    int nbv_lower;    /* s_nbv_xmin: rows of derivs_xmin */
    int nbv_upper;    /* s_nbv_xmax */
    int matrix_kind;  /* 0 dense, 1 band, 2 pds band, 3 pds tridiag (main block) */
    int blocked;      /* 0 plain, 1 2x2 blocks, 2 3x3 blocks */
    int main_size;    /* size of the band main block */
    int border;       /* nb = n - main_size */
    double dx;        /* m_dx, average cell size */
    int kernel;       /* build kernel used: 0 generic (one thread per line), 1 tiled (register-resident),
                         2 tiled FP32 sweeps with FP64 refinement (precision DDCX_FP32_REFINE) */
    int pivoted;      /* 1 if the band factorisation of the main block interchanged rows */
    int refine_steps; /* kernel 2: refinement steps after the first FP32 solve (1 or 2), else 0 */
    double refine_contraction; /* kernel 2: relative error of one FP32 solve, measured at create */
} ddcx_builder_info;

ddcx_status ddcx_bsplines_create(const ddcx_bsplines_desc* desc, ddcx_bsplines** out);
ddcx_status ddcx_bsplines_destroy(ddcx_bsplines* b);
/* knots: host output of ncells+1+2*degree coordinates (ddc::coordinate of the knot domain). */
ddcx_status ddcx_bsplines_knots(const ddcx_bsplines* b, double* knots);
/* ddc::integrals(), integrals.hpp:165-177: host output of size() values. */
ddcx_status ddcx_bsplines_integrals(const ddcx_bsplines* b, double* out);

ddcx_status ddcx_builder_create(const ddcx_bsplines* b, const ddcx_builder_desc* desc, ddcx_builder** out);
ddcx_status ddcx_builder_destroy(ddcx_builder* h);
ddcx_status ddcx_builder_get_info(const ddcx_builder* h, ddcx_builder_info* info);
/* Interpolation points (host output, npoints values). */
ddcx_status ddcx_builder_points(const ddcx_builder* h, double* points);

/* SplineBuilder::operator()(spline, vals, derivs_xmin, derivs_xmax),
 * spline_builder.hpp:538-557 / :896-1101.  coef: size() rows per line;
 * vals: npoints rows; derivs_*: nbv rows (NULL unless the closure is HERMITE),
 * both described by derivs_layout. */
ddcx_status ddcx_builder_build(const ddcx_builder* h, double* coef, ddcx_layout coef_layout, const double* vals,
                               ddcx_layout vals_layout, const double* derivs_xmin, const double* derivs_xmax,
                               ddcx_layout derivs_layout, int64_t nbatch, void* stream);

/* SplineBuilder::operator() then SplineEvaluator::operator() on the same lines
 * (spline_builder.hpp:905, spline_evaluator.hpp:263) in one pass (no reference
 * counterpart as one call).  The evaluator must use the builder's basis.  coef
 * may be NULL: the coefficients are then kept on chip only.  The fused kernel
 * covers FP64 builders with a tiled plan, degrees 2..5, and batch-contiguous
 * vals, coef and points; other cases run the two calls and need coef. */
ddcx_status ddcx_build_eval(const ddcx_builder* h, const ddcx_evaluator* e, double* coef, ddcx_layout coef_layout,
                            const double* vals, ddcx_layout vals_layout, const double* derivs_xmin,
                            const double* derivs_xmax, ddcx_layout derivs_layout, double* out, const double* coords,
                            ddcx_layout points_layout, int64_t npts, int64_t nbatch, void* stream);

/* One step of the semi-Lagrangian advection of the reference's example
 * (examples/characteristics_advection.cpp:237-274, the "feet" of
 * benchmarks/splines.cpp:188-206): the spline of vals is built along the
 * interpolation dimension and evaluated at the feet x_i - shift[l] of every
 * interpolation point x_i of line l, all in one fused kernel.  The feet are
 * computed in the kernel and the coefficients stay on chip; coef may be NULL,
 * and when it is given the coefficients are stored too.  out has npoints rows
 * per line, described by out_layout.  Covers what ddcx_build_eval's fused
 * kernel covers; other cases return DDCX_ERR_UNSUPPORTED. */
ddcx_status ddcx_advect(const ddcx_builder* h, const ddcx_evaluator* e, double* coef, ddcx_layout coef_layout,
                        const double* vals, ddcx_layout vals_layout, const double* derivs_xmin,
                        const double* derivs_xmax, ddcx_layout derivs_layout, const double* shift, double* out,
                        ddcx_layout out_layout, int64_t nbatch, void* stream);

/* Vector-Jacobian product of ddcx_builder_build (no reference counterpart;
 * built on the transposed solve SplinesLinearProblem::solve(b, true),
 * splines_linear_problem.hpp:77): given gcoef, writes gvals and the Hermite
 * derivative gradients (NULL when not Hermite). */
ddcx_status ddcx_builder_build_vjp(const ddcx_builder* h, const double* gcoef, ddcx_layout coef_layout,
                                   double* gvals, ddcx_layout vals_layout, double* gderivs_xmin,
                                   double* gderivs_xmax, ddcx_layout derivs_layout, int64_t nbatch, void* stream);

/* SplineBuilder::quadrature_coefficients(), spline_builder.hpp:1111-1234.
 * Host outputs: derivs_xmin (nbv_lower), values (nbasis - nbv_lower - nbv_upper;
 * this equals npoints except for HOMOGENEOUS_HERMITE closures, where the
 * reference also returns the weights of the zero derivative rows),
 * derivs_xmax (nbv_upper). */
ddcx_status ddcx_builder_quadrature(const ddcx_builder* h, double* q_xmin, double* q_vals, double* q_xmax);

/* SplineEvaluator<Exec, Mem, BSplines, EvalDDim, LowerRule, UpperRule> ctor,
 * spline_evaluator.hpp:140-153.  A Constant rule made here evaluates at the
 * domain bound it guards (rmin for the lower rule, rmax for the upper), i.e.
 * ConstantExtrapolationRule(rmin) / (rmax); in a 2D evaluator it clamps the
 * other coordinate to that dimension's [rmin, rmax]. */
ddcx_status ddcx_evaluator_create(const ddcx_bsplines* b, ddcx_extrap lower, ddcx_extrap upper,
                                  ddcx_evaluator** out);

/* An extrapolation rule with its constructor arguments.  For kind
 * DDCX_EXTRAP_CONSTANT:
 *  - eval_pos: ConstantExtrapolationRule<DimI>(eval_pos)
 *    (constant_extrapolation_rule.hpp:39-40): out-of-domain points take the
 *    spline's value (and, in the adjoint, basis weights) at eval_pos, which
 *    must lie in [rmin, rmax];
 *  - eval_pos_not_interest_min/max: the 2D form ConstantExtrapolationRule<DimI,
 *    DimNI>(eval_pos, min, max) (:99-124): in ddcx_evaluator2d_* calls, when the
 *    other dimension is not periodic, its coordinate is clamped to [min, max].
 * NaN selects the default: the guarded bound for eval_pos, the other
 * dimension's [rmin, rmax] for the pair (give both or neither).  Other kinds
 * ignore the positions. */
typedef struct ddcx_extrap_rule {
    ddcx_extrap kind;
    double eval_pos;
    double eval_pos_not_interest_min;
    double eval_pos_not_interest_max;
} ddcx_extrap_rule;

/* SplineEvaluator ctor with rule objects (spline_evaluator.hpp:140-153,
 * taking LowerExtrapolationRule / UpperExtrapolationRule by value).
 * DDCX_ERR_INVALID for an eval_pos outside [rmin, rmax]. */
ddcx_status ddcx_evaluator_create_with_rules(const ddcx_bsplines* b, const ddcx_extrap_rule* lower,
                                             const ddcx_extrap_rule* upper, ddcx_evaluator** out);
ddcx_status ddcx_evaluator_destroy(ddcx_evaluator* h);

/* SplineEvaluator::operator()(spline_eval, coords_eval, spline_coef),
 * spline_evaluator.hpp:257-294: out and coords share points_layout (npts rows). */
ddcx_status ddcx_evaluator_eval(const ddcx_evaluator* h, double* out, const double* coords,
                                ddcx_layout points_layout, int64_t npts, const double* coef,
                                ddcx_layout coef_layout, int64_t nbatch, void* stream);

/* SplineEvaluator::operator()(spline_eval, spline_coef) and
 * deriv(order, spline_eval, spline_coef), spline_evaluator.hpp:296-342,
 * :432-478: the evaluation points are the evaluation domain's own
 * coordinates, the same for every line: `points` holds npts coordinates,
 * point p of every line at points[p]; out has points_layout.  order 0 is
 * operator() (with the rules), order > 0 deriv(). */
ddcx_status ddcx_evaluator_eval_at(const ddcx_evaluator* h, int order, double* out, ddcx_layout points_layout,
                                   const double* points, int64_t npts, const double* coef, ddcx_layout coef_layout,
                                   int64_t nbatch, void* stream);

/* SplineEvaluator::deriv(DiscreteElement<Deriv<X>> order, ...), spline_evaluator.hpp:357-430. */
ddcx_status ddcx_evaluator_deriv(const ddcx_evaluator* h, int order, double* out, const double* coords,
                                 ddcx_layout points_layout, int64_t npts, const double* coef,
                                 ddcx_layout coef_layout, int64_t nbatch, void* stream);

/* The device data of an evaluator, for the device-callable single-point
 * evaluation of include/ddcx_device.hpp (SplineEvaluator::operator()(coord,
 * coef_1d) and deriv(order, coord, coef_1d), spline_evaluator.hpp:228-235,
 * :357-478): the basis description (knots and cell brackets in device
 * memory of the evaluator's basis, valid while the evaluator exists) and the
 * extrapolation rules with their positions. */
typedef struct ddcx_device_evaluator {
    int degree, uniform, periodic, ncells;
    double origin, step, inv_step; /* uniform knots: origin + k * step */
    double rmin, rmax, length;
    const double* knots;   /* non-uniform knots (ncells + 1 + 2 * degree), device memory */
    const uint32_t* cells; /* non-uniform cell-search brackets, device memory (NULL: full search) */
    int nbucket;
    double inv_bucket;
    int lower, upper;      /* ddcx_extrap of the lower and upper rule */
    double pos_lower, pos_upper; /* evaluation positions of Constant rules */
} ddcx_device_evaluator;
ddcx_status ddcx_evaluator_device_view(const ddcx_evaluator* h, ddcx_device_evaluator* out);

/* SplineEvaluator::integrate(integrals, spline_coef), spline_evaluator.hpp:493-526. */
ddcx_status ddcx_evaluator_integrate(const ddcx_evaluator* h, double* out, int64_t out_stride, const double* coef,
                                     ddcx_layout coef_layout, int64_t nbatch, void* stream);

/* Vector-Jacobian product of ddcx_evaluator_eval (no reference counterpart):
 * gcoef = sum_p w_p dB(x_p)/dcoef (segmented scatter-add, overwritten),
 * gx_p = w_p f'(x_p) (deriv() order 1; 0 where an extrapolation rule applies). */
ddcx_status ddcx_evaluator_eval_vjp(const ddcx_evaluator* h, const double* coords, const double* w, double* gx,
                                    ddcx_layout points_layout, int64_t npts, const double* coef, double* gcoef,
                                    ddcx_layout coef_layout, int64_t nbatch, void* stream);

/* SplineEvaluator2D::operator() / deriv*(), spline_evaluator_2d.hpp:370-600
 * (eval_no_bc :740-825): batched 2D evaluation.  coords: interleaved (x1, x2)
 * pairs, point p of batch l at coords[2*(l*npts + p)], out[l*npts + p]; coef of
 * batch l at coef + l*coef_batch_stride, element (i1, i2) at i1*st1 + i2*st2.
 * rules: {lower1, upper1, lower2, upper2} extrapolation rules. */
ddcx_status ddcx_evaluator2d_eval(const ddcx_evaluator* e1, const ddcx_evaluator* e2, int order1, int order2,
                                  double* out, const double* coords, int64_t npts, const double* coef,
                                  int64_t coef_batch_stride, int64_t st1, int64_t st2, int64_t nbatch,
                                  void* stream);

/* ddcx_evaluator2d_eval through a packed copy of the coefficients (no
 * reference counterpart; the same values to rounding): for many points per
 * batch entry, each cell's (d1+1) x (d2+1) coefficient patch is first copied
 * into one contiguous record in `work`, so that a point gathers one record
 * instead of d1+1 separate rows.  work: device memory, 16-byte aligned, of at
 * least the bytes ddcx_evaluator2d_eval_workspace gives for one batch entry;
 * with less than all of nbatch, the batch entries are packed and evaluated a
 * workspace-full at a time.  Calls that share a workspace must not overlap. */
ddcx_status ddcx_evaluator2d_eval_workspace(const ddcx_evaluator* e1, const ddcx_evaluator* e2, int64_t nbatch,
                                            int64_t* bytes);
ddcx_status ddcx_evaluator2d_eval_packed(const ddcx_evaluator* e1, const ddcx_evaluator* e2, int order1, int order2,
                                         double* out, const double* coords, int64_t npts, const double* coef,
                                         int64_t coef_batch_stride, int64_t st1, int64_t st2, int64_t nbatch,
                                         void* work, int64_t work_bytes, void* stream);

/* ddcx_evaluator2d_eval_vjp with the points binned by cell tile (no
 * reference counterpart): the same results, with per-tile accumulation in
 * LDS instead of one global atomic per point and coefficient.  work: device
 * memory of at least *bytes as given by ddcx_evaluator2d_eval_vjp_workspace
 * (0 when the grid or degrees are out of the binned kernels' range: then use
 * ddcx_evaluator2d_eval_vjp).  Calls that share a workspace must not overlap. */
ddcx_status ddcx_evaluator2d_eval_vjp_workspace(const ddcx_evaluator* e1, const ddcx_evaluator* e2, int64_t npts,
                                                int64_t nbatch, int64_t* bytes);
ddcx_status ddcx_evaluator2d_eval_vjp_binned(const ddcx_evaluator* e1, const ddcx_evaluator* e2,
                                             const double* coords, const double* w, double* gx, int64_t npts,
                                             const double* coef, double* gcoef, int64_t coef_batch_stride,
                                             int64_t st1, int64_t st2, int64_t nbatch, void* work,
                                             int64_t work_bytes, void* stream);

/* ddcx_evaluator2d_eval_vjp_binned without a workspace argument: no
 * reference counterpart; for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The workspace is a static device array of the library,
 * DDCX_STATIC_SCRATCH_BYTES per GPU (about 30 B per point: up to ~2.2M points
 * per call), held by every GPU the library runs on until the process ends;
 * DDCX_ERR_INVALID when the call needs more.  Nothing is allocated.  All
 * calls on one GPU share that array: calls that overlap on one GPU (other
 * streams, other threads) race and must be serialised by the caller. */
#define DDCX_STATIC_SCRATCH_BYTES (64u << 20)
ddcx_status ddcx_evaluator2d_eval_vjp_binned_static(const ddcx_evaluator* e1, const ddcx_evaluator* e2,
                                                    const double* coords, const double* w, double* gx, int64_t npts,
                                                    const double* coef, double* gcoef, int64_t coef_batch_stride,
                                                    int64_t st1, int64_t st2, int64_t nbatch, void* stream);

/* SplineEvaluatorND::operator() / deriv(), spline_evaluator_nd.hpp:300-590
 * (eval :684-720, eval_no_bc :782-826); SplineEvaluator3D
 * (spline_evaluator_3d.hpp:443-1019) is its nd = 3 case.  evs: one evaluator
 * per dimension (1 <= nd <= 4); their extrapolation rules must be Null or
 * Periodic (Constant is refused: its reference form exists for 1D and 2D
 * only).  orders: nd derivative orders (NULL: all 0), each at most the
 * degree (spline_evaluator_nd.hpp:735); with any nonzero order no wrap or rule is applied, as in
 * deriv().  coords: the nd coordinates of point p of batch l at
 * coords[nd*(l*npts + p)], out[l*npts + p]; coef of batch l at
 * coef + l*coef_batch_stride, element (i_1..i_nd) at sum_k i_k*coef_strides[k]. */
ddcx_status ddcx_evaluatornd_eval(const ddcx_evaluator* const* evs, int nd, const int* orders, double* out,
                                  const double* coords, int64_t npts, const double* coef, int64_t coef_batch_stride,
                                  const int64_t* coef_strides, int64_t nbatch, void* stream);

/* ddcx_evaluatornd_eval with the points binned by tile of cells (no
 * reference counterpart; the same values bit for bit): for three dimensions
 * of one degree and many scattered points per batch entry, each tile's
 * coefficient patch is read once into on-chip memory instead of every point
 * gathering (d+1)^2 rows of the coefficient array.  *bytes = the device
 * workspace for npts x nbatch points (about 34 B per point), 0 when the grid
 * is outside the binned kernels' range (then use ddcx_evaluatornd_eval).
 * work: 8-byte aligned; calls that share a workspace must not overlap. */
ddcx_status ddcx_evaluatornd_eval_workspace(const ddcx_evaluator* const* evs, int nd, int64_t npts, int64_t nbatch,
                                            int64_t* bytes);
ddcx_status ddcx_evaluatornd_eval_binned(const ddcx_evaluator* const* evs, int nd, const int* orders, double* out,
                                         const double* coords, int64_t npts, const double* coef,
                                         int64_t coef_batch_stride, const int64_t* coef_strides, int64_t nbatch,
                                         void* work, int64_t work_bytes, void* stream);

/* SplineEvaluatorND::integrate(integrals, spline_coef),
 * spline_evaluator_nd.hpp:597-632: out[l*out_stride] = integral of batch l's
 * spline over the whole domain. */
ddcx_status ddcx_evaluatornd_integrate(const ddcx_evaluator* const* evs, int nd, double* out, int64_t out_stride,
                                       const double* coef, int64_t coef_batch_stride, const int64_t* coef_strides,
                                       int64_t nbatch, void* stream);

/* The same integral with a caller workspace (no reference counterpart): the
 * rows of each batch entry are split over several workgroups whose partial
 * sums go to the workspace and are then added in index order, so a small
 * batch still spreads over the whole GPU.  *bytes = the workspace size for
 * nbatch entries (0 when the split is not needed; work may then be NULL).
 * Results equal ddcx_evaluatornd_integrate's up to rounding. */
ddcx_status ddcx_evaluatornd_integrate_workspace(const ddcx_evaluator* const* evs, int nd, int64_t nbatch,
                                                 int64_t* bytes);
ddcx_status ddcx_evaluatornd_integrate_ws(const ddcx_evaluator* const* evs, int nd, double* out, int64_t out_stride,
                                          const double* coef, int64_t coef_batch_stride, const int64_t* coef_strides,
                                          int64_t nbatch, void* work, int64_t work_bytes, void* stream);

/* Vector-Jacobian product of ddcx_evaluator2d_eval (value form; no reference
 * counterpart): gcoef (same strides as coef, overwritten) = sum_p w_p
 * B1(x1_p) (x) B2(x2_p) (tensor scatter-add); gx (interleaved pairs like
 * coords) = w_p (df/dx1, df/dx2), 0 for a coordinate clamped by the constant
 * rule and for both under the Null rule. */
ddcx_status ddcx_evaluator2d_eval_vjp(const ddcx_evaluator* e1, const ddcx_evaluator* e2, const double* coords,
                                      const double* w, double* gx, int64_t npts, const double* coef, double* gcoef,
                                      int64_t coef_batch_stride, int64_t st1, int64_t st2, int64_t nbatch,
                                      void* stream);

/* ---- The linear problem on its own ------------------------------------------
 * SplinesLinearProblem<ExecSpace> (splines_linear_problem.hpp:26-127): the
 * matrix object the builder owns, as made by SplinesLinearProblemMaker
 * (splines_linear_problem_maker.hpp:25-170).  Elements are set on the host
 * before setup; setup factors on the host in FP64 (the LAPACK algorithms of
 * the reference) and uploads the factors; solves then run on the device with
 * the sweep code of the builds. */
typedef enum ddcx_problem_kind {
    DDCX_PROBLEM_DENSE = 0,         /* SplinesLinearProblemDense(n) */
    DDCX_PROBLEM_BAND = 1,          /* make_new_band(n, kl, ku, pds) */
    DDCX_PROBLEM_BLOCK_BAND = 2,    /* make_new_block_matrix_with_band_main_block(n, kl, ku, pds, bottom, top) */
    DDCX_PROBLEM_PERIODIC_BAND = 3, /* make_new_periodic_band_matrix(n, kl, ku, pds) */
    DDCX_PROBLEM_BLOCK_DENSE = 4,   /* SplinesLinearProblem2x2Blocks(n, Dense(n - bottom)) for top == 0, else
                                       SplinesLinearProblem3x3Blocks(n, top, Dense(n - top - bottom)) */
    DDCX_PROBLEM_BAND_STORAGE = 5   /* SplinesLinearProblemBand(n, kl, ku) built directly (the maker picks
                                       dense storage for wide bands) */
} ddcx_problem_kind;

typedef struct ddcx_problem_desc {
    ddcx_problem_kind kind;
    int n;           /* matrix size */
    int kl, ku;      /* band kinds: sub- and super-diagonals */
    int pds;         /* band kinds: positive-definite symmetric band */
    int bottom_size; /* block kinds: bottom-right block size */
    int top_size;    /* block kinds: top-left block size (3x3 blocks when > 0) */
} ddcx_problem_desc;

typedef struct ddcx_problem_info {
    int size;        /* SplinesLinearProblem::size() */
    int rhs_rows;    /* required_number_of_rhs_rows() */
    int main_kind;   /* 0 dense, 1 band, 2 pds band, 3 pds tridiag (the main block, or the whole matrix) */
    int blocked;     /* 0 plain, 1 2x2 blocks, 2 3x3 blocks */
    int main_size;
    int pivoted;     /* after setup: 1 if a band or dense factor interchanged rows */
} ddcx_problem_info;

typedef struct ddcx_problem ddcx_problem;

ddcx_status ddcx_problem_create(const ddcx_problem_desc* desc, ddcx_problem** out);
ddcx_status ddcx_problem_destroy(ddcx_problem* h);
/* set_element / get_element (splines_linear_problem.hpp:63-70); set only
 * before setup.  After setup, get returns the assembled (unfactored) entry. */
ddcx_status ddcx_problem_set_element(ddcx_problem* h, int i, int j, double aij);
ddcx_status ddcx_problem_get_element(const ddcx_problem* h, int i, int j, double* aij);
/* setup_solver() (splines_linear_problem.hpp:72): factor and upload.
 * DDCX_ERR_SINGULAR with the LAPACK info as the reference throws it. */
ddcx_status ddcx_problem_setup_solver(ddcx_problem* h);
ddcx_status ddcx_problem_get_info(const ddcx_problem* h, ddcx_problem_info* info);
/* solve(b, transpose) (splines_linear_problem.hpp:77), in place on the device:
 * b holds nrhs right-hand sides, row i of right-hand side j at b[i * ldb + j]
 * (the reference's LayoutRight MultiRHS view (rhs_rows, nrhs)); rows past
 * size() are workspace, as in the reference.  Stream-ordered. */
ddcx_status ddcx_problem_solve(const ddcx_problem* h, double* b, int64_t ldb, int64_t nrhs, int transpose,
                               void* stream);

const char* ddcx_last_error(void);
const char* ddcx_version(void);

#ifdef __cplusplus
}
#endif

#endif /* DDCX_H */
