// B-spline basis kernels shared by the host setup (matrix assembly) and the
// device evaluators.  Every routine is a restatement of the reference's
// UniformBSplines / NonUniformBSplines Impl methods and keeps their operation
// order, so that cell indices are bit-identical and values match to rounding.
// The whole library is compiled with -ffp-contract=off: no multiply-add is
// fused, exactly as in a plain FP64 CPU build of the reference.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ddcx {

constexpr int kMaxDegree = 7;

// Value-type description of a B-spline basis, passed by value to kernels.
struct BSplinesView {
    int degree;
    int uniform;
    int periodic;
    int ncells;
    // uniform knots: coordinate(k) = origin + k * step
    // (UniformPointSampling::init_ghosted, uniform_point_sampling.hpp:163-192)
    double origin;
    double step;
    double inv_step; // 1.0 / step, bsplines_uniform.hpp:357-360
    double rmin;     // coordinate(break_point_domain.front())
    double rmax;     // coordinate(break_point_domain.back())
    double length;   // rmax - rmin
    const double* knots; // non-uniform knots (ncells + 1 + 2*degree), device or host memory
    // Non-uniform cell search brackets: nbucket equal buckets over [rmin, rmax];
    // cells[k] packs the first (low 16 bits) and last (high 16 bits) cell that
    // bucket k overlaps.  nullptr: plain binary search over all cells.
    const uint32_t* cells;
    int nbucket;
    double inv_bucket; // nbucket / (rmax - rmin)
};

__host__ __device__ inline double knot_coord(BSplinesView const& b, int k)
{
    if (b.uniform) return b.origin + static_cast<double>(k) * b.step;
    return b.knots[k];
}

// UniformBSplines::Impl::get_icell_and_offset, bsplines_uniform.hpp:568-595
__host__ __device__ inline int uniform_cell(BSplinesView const& b, double x, double& offset)
{
    int icell;
    if (x <= b.rmin) {
        icell = 0;
        offset = 0.0;
    } else if (x >= b.rmax) {
        icell = b.ncells - 1;
        offset = 1.0;
    } else {
        offset = (x - b.rmin) * b.inv_step;
        icell = static_cast<int>(offset);
        offset = offset - icell;
        if (icell == b.ncells && offset == 0.0) {
            icell = b.ncells - 1;
            offset = 1.0;
        }
    }
    return icell;
}

// NonUniformBSplines::Impl::find_cell_start, bsplines_non_uniform.hpp:668-695.
// Returns the knot index of the cell start (break point domain starts at
// knot index `degree`).
__host__ __device__ inline int nonuniform_cell(BSplinesView const& b, const double* knots, double x)
{
    int const front = b.degree;
    int const back = b.degree + b.ncells;
    if (x <= knots[front]) return front;
    if (x >= knots[back]) return back - 1;
    int low = front;
    int high = back;
    if (b.cells) {
        // Start the same search from the bracket of x's bucket.  The result is
        // the unique cell with knots[icell] <= x < knots[icell + 1], so it is
        // the reference's whatever bracket holds it; a bracket missed through
        // rounding of the bucket index falls back to the full range.
        int k = int((x - b.rmin) * b.inv_bucket);
        k = k < 0 ? 0 : (k >= b.nbucket ? b.nbucket - 1 : k);
        uint32_t const e = b.cells[k];
        int const lo = front + int(e & 0xffffu);
        int const hi = front + int(e >> 16) + 1;
        if (x >= knots[lo]) low = lo;
        if (x < knots[hi]) high = hi;
    }
    int icell = low + (high - low) / 2;
    while (x < knots[icell] || x >= knots[icell + 1]) {
        if (x < knots[icell]) high = icell;
        else low = icell;
        icell = low + (high - low) / 2;
    }
    return icell;
}

// The extrapolation rules of a 1D evaluator (SplineEvaluator's lower and
// upper rules, spline_evaluator.hpp:140-153): kind 0 Null, 1 Constant.  A
// Constant rule evaluates the spline at its eval_pos instead of the point
// (ConstantExtrapolationRule::m_eval_pos, constant_extrapolation_rule.hpp:39-72).
struct Rules1D {
    int lower, upper;
    double pos_lower, pos_upper;
};

// The four rules of a 2D evaluator, lower and upper of dimension 1 then of
// dimension 2 (SplineEvaluator2D, spline_evaluator_2d.hpp:679-720).  A
// Constant rule of dimension k evaluates at pos in dimension k and, when the
// other dimension is not periodic, clamps the other coordinate to
// [ni_min, ni_max] (constant_extrapolation_rule.hpp:99-124, :160-172).
struct Rules2D {
    int kind[4];
    double pos[4];
    double ni_min[4], ni_max[4];
};

// SplineEvaluator::eval, spline_evaluator.hpp:529-554: periodic wrap or
// extrapolation.  Returns false when the Null rule gives 0; sets x to the
// position actually evaluated (the rule's eval_pos, for the Constant rule)
// and `outside` when a rule applied.
__host__ __device__ inline bool apply_bc(BSplinesView const& b, Rules1D const& r, double& x, bool& outside)
{
    outside = false;
    if (b.periodic) {
        if (x < b.rmin || x > b.rmax) x -= floor((x - b.rmin) / b.length) * b.length;
        return true;
    }
    if (x < b.rmin) {
        outside = true;
        if (r.lower == 0) return false;
        x = r.pos_lower;
    } else if (x > b.rmax) {
        outside = true;
        if (r.upper == 0) return false;
        x = r.pos_upper;
    }
    return true;
}

// SplineEvaluator2D::eval, spline_evaluator_2d.hpp:679-720: periodic wraps,
// then the first rule that applies (dimension 1 before dimension 2).  Returns
// false when a Null rule gives 0; otherwise (x1, x2) is the position
// evaluated, and in1 / in2 say whether that coordinate still depends on the
// point's (false where a rule replaced or clamped it).
__host__ __device__ inline bool place2d(BSplinesView const& b1, BSplinesView const& b2, Rules2D const& r, double& x1,
                                        double& x2, bool& in1, bool& in2)
{
    in1 = in2 = true;
    if (b1.periodic && (x1 < b1.rmin || x1 > b1.rmax)) x1 -= floor((x1 - b1.rmin) / b1.length) * b1.length;
    if (b2.periodic && (x2 < b2.rmin || x2 > b2.rmax)) x2 -= floor((x2 - b2.rmin) / b2.length) * b2.length;
    auto clamp = [](double v, double lo, double hi, bool& in) {
        if (v < lo) {
            in = false;
            return lo;
        }
        if (v > hi) {
            in = false;
            return hi;
        }
        return v;
    };
    if (!b1.periodic && (x1 < b1.rmin || x1 > b1.rmax)) {
        int const k = x1 < b1.rmin ? 0 : 1;
        if (r.kind[k] == 0) return false;
        x1 = r.pos[k];
        in1 = false;
        if (!b2.periodic) x2 = clamp(x2, r.ni_min[k], r.ni_max[k], in2);
        return true;
    }
    if (!b2.periodic && (x2 < b2.rmin || x2 > b2.rmax)) {
        int const k = x2 < b2.rmin ? 2 : 3;
        if (r.kind[k] == 0) return false;
        x2 = r.pos[k];
        in2 = false;
        if (!b1.periodic) x1 = clamp(x1, r.ni_min[k], r.ni_max[k], in1);
    }
    return true;
}

// 1/d for the evaluation kernels: the hardware reciprocal refined by two
// Newton-Raphson steps, r <- r + r (1 - d r), within a couple of ulps of the
// correctly rounded quotient.  Cell indices never depend on it; the host
// setup (matrix assembly, integrals) keeps exact divisions.
__device__ __forceinline__ double fast_rcp(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    return fma(r, e, r);
}

// Cox-de Boor on the unit grid, UniformBSplines eval_basis,
// bsplines_uniform.hpp:394-426.  NV values (NV = degree+1; integrals() uses
// degree+2).
// FAST (device evaluation only): multiplications by the constants 1/j
// instead of divisions by j.
// FAST also fuses the update saved + xx * temp into one fma.
template <int NV, bool FAST = false>
__host__ __device__ inline void uniform_basis_from_offset(double offset, double* values)
{
    values[0] = 1.0;
#pragma unroll
    for (int j = 1; j < NV; ++j) {
        double xx = -offset;
        double saved = 0.0;
#pragma unroll
        for (int r = 0; r < j; ++r) {
            xx += 1;
            double const temp = FAST ? values[r] * (1.0 / j) : values[r] / j;
            values[r] = FAST ? fma(xx, temp, saved) : saved + xx * temp;
            saved = (j - xx) * temp;
        }
        values[j] = saved;
    }
}

// de Boor with left/right arrays, NonUniformBSplines eval_basis,
// bsplines_non_uniform.hpp:461-497; `icell` is the knot index from
// nonuniform_cell.  FAST (device evaluation): the update is one fma, and the
// D(D+1)/2 divisors right[r] + left[j - r], which do not depend on the
// recurrence's values, are inverted together: one refined hardware
// reciprocal of their product, then each inverse as that reciprocal times
// the product of the others (prefix products forward, the running inverse
// backward), 3 multiplications per divisor instead of a reciprocal and four
// fmas each. Every divisor is at least the cell's width, so the product is
// positive; outside [2^-960, 2^960] each divisor takes its own reciprocal.
template <int D, bool FAST = false>
__host__ __device__ inline void nonuniform_basis(const double* knots, int icell, double x, double* values)
{
    double left[D > 0 ? D : 1];
    double right[D > 0 ? D : 1];
    values[0] = 1.0;
    if constexpr (FAST && D > 0) {
        constexpr int M = D * (D + 1) / 2;
#pragma unroll
        for (int j = 0; j < D; ++j) {
            left[j] = x - knots[icell - j];
            right[j] = knots[icell + j + 1] - x;
        }
        double den[M], pre[M], inv[M];
        {
            int k = 0;
#pragma unroll
            for (int j = 0; j < D; ++j)
#pragma unroll
                for (int r = 0; r < j + 1; ++r) den[k++] = right[r] + left[j - r];
        }
        pre[0] = den[0];
#pragma unroll
        for (int k = 1; k < M; ++k) pre[k] = pre[k - 1] * den[k];
        double const mag = fabs(pre[M - 1]);
        if (mag >= 0x1p-960 && mag <= 0x1p960) {
            double rv = fast_rcp(pre[M - 1]);
#pragma unroll
            for (int k = M - 1; k > 0; --k) {
                inv[k] = rv * pre[k - 1];
                rv *= den[k];
            }
            inv[0] = rv;
        } else {
#pragma unroll
            for (int k = 0; k < M; ++k) inv[k] = fast_rcp(den[k]);
        }
        int k = 0;
#pragma unroll
        for (int j = 0; j < D; ++j) {
            double saved = 0.0;
#pragma unroll
            for (int r = 0; r < j + 1; ++r) {
                double const temp = values[r] * inv[k++];
                values[r] = fma(right[r], temp, saved);
                saved = left[j - r] * temp;
            }
            values[j + 1] = saved;
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < D; ++j) {
        left[j] = x - knots[icell - j];
        right[j] = knots[icell + j + 1] - x;
        double saved = 0.0;
#pragma unroll
        for (int r = 0; r < j + 1; ++r) {
            double const temp = values[r] / (right[r] + left[j - r]);
            values[r] = saved + right[r] * temp;
            saved = left[j - r] * temp;
        }
        values[j + 1] = saved;
    }
}

// Basis kind of a kernel instantiation: decided by BSplinesView::uniform at
// run time, or fixed at compile time so that the other branch is not emitted.
enum BasisKind : int { kAnyBasis = -1, kNonUniformBasis = 0, kUniformBasis = 1 };

__host__ __device__ constexpr bool is_uniform_kind(int kind, int uniform)
{
    return kind == kAnyBasis ? uniform != 0 : kind == kUniformBasis;
}

// eval_basis for either basis kind; returns jmin (first B-spline index).
template <int D, bool FAST = false, int KIND = kAnyBasis>
__host__ __device__ inline int eval_basis(BSplinesView const& b, const double* knots, double x, double* values)
{
    if (is_uniform_kind(KIND, b.uniform)) {
        double offset;
        int const jmin = uniform_cell(b, x, offset);
        uniform_basis_from_offset<D + 1, FAST>(offset, values);
        return jmin;
    }
    int const icell = nonuniform_cell(b, knots, x);
    nonuniform_basis<D, FAST>(knots, icell, x, values);
    return icell - D;
}

// First B-spline index of the cell holding x: the jmin that eval_basis and
// eval_basis_and_n_derivs return for the same x, without the basis values.
__host__ __device__ inline int cell_jmin(BSplinesView const& b, double x)
{
    if (b.uniform) {
        double offset;
        return uniform_cell(b, x, offset);
    }
    return nonuniform_cell(b, b.knots, x) - b.degree;
}

// eval_basis_and_n_derivs (NURBS book A2.3), uniform bsplines_uniform.hpp:475-564
// and non-uniform bsplines_non_uniform.hpp:561-664.  derivs is (D+1) x (N+1)
// row-major.  Returns jmin.
template <int D, int N, bool FAST = false, int KIND = kAnyBasis>
__host__ __device__ inline int eval_basis_and_n_derivs(BSplinesView const& b, const double* knots, double x,
                                                       double* derivs)
{
    double ndu[D + 1][D + 1];
    double a[D + 1][2];
    int jmin;
    bool const uni = is_uniform_kind(KIND, b.uniform);
    if (uni) {
        double offset;
        jmin = uniform_cell(b, x, offset);
        ndu[0][0] = 1.0;
#pragma unroll
        for (int j = 1; j < D + 1; ++j) {
            double xx = -offset;
            double saved = 0.0;
#pragma unroll
            for (int r = 0; r < j; ++r) {
                xx += 1.0;
                double const temp = FAST ? ndu[j - 1][r] * (1.0 / j) : ndu[j - 1][r] / j;
                ndu[j][r] = FAST ? fma(xx, temp, saved) : saved + xx * temp;
                saved = (j - xx) * temp;
            }
            ndu[j][j] = saved;
        }
    } else {
        double left[D > 0 ? D : 1];
        double right[D > 0 ? D : 1];
        int const icell = nonuniform_cell(b, knots, x);
        jmin = icell - D;
        ndu[0][0] = 1.0;
#pragma unroll
        for (int j = 0; j < D; ++j) {
            left[j] = x - knots[icell - j];
            right[j] = knots[icell + j + 1] - x;
            double saved = 0.0;
#pragma unroll
            for (int r = 0; r < j + 1; ++r) {
                if constexpr (FAST) ndu[r][j + 1] = fast_rcp(right[r] + left[j - r]);
                else ndu[r][j + 1] = 1.0 / (right[r] + left[j - r]);
                double const temp = ndu[j][r] * ndu[r][j + 1];
                ndu[j + 1][r] = FAST ? fma(right[r], temp, saved) : saved + right[r] * temp;
                saved = left[j - r] * temp;
            }
            ndu[j + 1][j + 1] = saved;
        }
    }
#pragma unroll
    for (int i = 0; i < D + 1; ++i) derivs[i * (N + 1)] = ndu[D][i];
#pragma unroll
    for (int r = 0; r < D + 1; ++r) {
        int s1 = 0;
        int s2 = 1;
        a[0][0] = 1.0;
#pragma unroll
        for (int k = 1; k < N + 1; ++k) {
            double d = 0.0;
            int const rk = r - k;
            int const pk = D - k;
            if (r >= k) {
                a[0][s2] = uni ? (FAST ? a[0][s1] * (1.0 / (pk + 1)) : a[0][s1] / (pk + 1))
                               : a[0][s1] * ndu[rk][pk + 1];
                d = a[0][s2] * ndu[pk][rk];
            }
            int const j1 = rk > -1 ? 1 : (-rk);
            int const j2 = (r - 1) <= pk ? k : (D - r + 1);
#pragma unroll
            for (int j = 1; j < D + 1; ++j) {
                if (j >= j1 && j < j2) {
                    a[j][s2] = uni ? (FAST ? (a[j][s1] - a[j - 1][s1]) * (1.0 / (pk + 1))
                                          : (a[j][s1] - a[j - 1][s1]) / (pk + 1))
                                   : (a[j][s1] - a[j - 1][s1]) * ndu[rk + j][pk + 1];
                    d = FAST ? fma(a[j][s2], ndu[pk][rk + j], d) : d + a[j][s2] * ndu[pk][rk + j];
                }
            }
            if (r <= pk) {
                a[k][s2] = uni ? (FAST ? -a[k - 1][s1] * (1.0 / (pk + 1)) : -a[k - 1][s1] / (pk + 1))
                               : -a[k - 1][s1] * ndu[r][pk + 1];
                d = FAST ? fma(a[k][s2], ndu[pk][r], d) : d + a[k][s2] * ndu[pk][r];
            }
            derivs[r * (N + 1) + k] = d;
            int const t = s1;
            s1 = s2;
            s2 = t;
        }
    }
    if (uni) {
        double d = D * b.inv_step;
#pragma unroll
        for (int k = 1; k < N + 1; ++k) {
#pragma unroll
            for (int i = 0; i < D + 1; ++i) derivs[i * (N + 1) + k] *= d;
            d *= (D - k) * b.inv_step;
        }
    } else {
        int r = D;
#pragma unroll
        for (int k = 1; k < N + 1; ++k) {
#pragma unroll
            for (int i = 0; i < D + 1; ++i) derivs[i * (N + 1) + k] *= r;
            r *= D - k;
        }
    }
    return jmin;
}

// LDS copies of the knot vectors and cell brackets of the non-uniform
// dimensions of a 2D kernel, laid out knots 1 | knots 2 | brackets 1 |
// brackets 2 (a uniform dimension has none).
inline size_t staged_table_bytes(BSplinesView const& b)
{
    return b.uniform ? 0
                     : size_t(b.ncells + 1 + 2 * b.degree) * sizeof(double) + size_t(b.nbucket) * sizeof(uint32_t);
}

inline size_t staged_tables_bytes_2d(BSplinesView const& b1, BSplinesView const& b2)
{
    return staged_table_bytes(b1) + staged_table_bytes(b2);
}

// whether the tables of both dimensions can be staged within `budget` bytes
// (a non-uniform dimension needs its cell brackets)
inline bool staged_tables_fit_2d(BSplinesView const& b1, BSplinesView const& b2, size_t budget)
{
    return (b1.uniform || b1.nbucket > 0) && (b2.uniform || b2.nbucket > 0) &&
           staged_tables_bytes_2d(b1, b2) <= budget;
}

#ifdef __HIPCC__
// The workgroup's threads copy the tables to `lds` (8-byte aligned,
// staged_tables_bytes_2d long) and point b1 / b2 at the copies, for either
// basis kind, so that every knot and bracket read of the kernel is an LDS
// read; the caller synchronises the workgroup before the first use.
__device__ inline void stage_tables_2d(BSplinesView& b1, BSplinesView& b2, double* lds, int tid, int nthreads)
{
    int const nk1 = b1.uniform ? 0 : b1.ncells + 1 + 2 * b1.degree, nk2 = b2.uniform ? 0 : b2.ncells + 1 + 2 * b2.degree;
    int const nc1 = b1.uniform ? 0 : b1.nbucket, nc2 = b2.uniform ? 0 : b2.nbucket;
    double* const kn1 = lds;
    double* const kn2 = lds + nk1;
    uint32_t* const cl1 = reinterpret_cast<uint32_t*>(lds + nk1 + nk2);
    uint32_t* const cl2 = cl1 + nc1;
    for (int i = tid; i < nk1; i += nthreads) kn1[i] = b1.knots[i];
    for (int i = tid; i < nc1; i += nthreads) cl1[i] = b1.cells[i];
    for (int i = tid; i < nk2; i += nthreads) kn2[i] = b2.knots[i];
    for (int i = tid; i < nc2; i += nthreads) cl2[i] = b2.cells[i];
    b1.knots = kn1;
    b1.cells = cl1;
    b2.knots = kn2;
    b2.cells = cl2;
}
#endif

// Dispatch a runtime degree to a compile-time one (host setup paths).
template <class F>
inline bool dispatch_degree(int degree, F&& f)
{
    switch (degree) {
    case 1: f.template operator()<1>(); return true;
    case 2: f.template operator()<2>(); return true;
    case 3: f.template operator()<3>(); return true;
    case 4: f.template operator()<4>(); return true;
    case 5: f.template operator()<5>(); return true;
    case 6: f.template operator()<6>(); return true;
    case 7: f.template operator()<7>(); return true;
    default: return false;
    }
}

} // namespace ddcx
