/*
 * ddc_oracle.c -- CPU restatement of DDC's batched spline path, used ONLY as a
 * test checker and as the timed CPU baseline.  It is never linked into, loaded
 * by, or called from the product library (ddc_amd/libddcx.so).
 *
 * Every function follows the reference algorithm step by step, in FP64, built
 * with -ffp-contract=off so that no multiply-add is fused (the reference's CPU
 * build does not contract either).  File:line citations are relative to the
 * DDC v0.15.1 source tree.
 *
 * Third-party routines the reference calls (LAPACKE dgbtrf/dpbtrf/dpttrf/
 * dgetrf, KokkosBatched Serial{Gbtrs,Pbtrs,Pttrs,Getrs}; kokkos-kernels 4.7,
 * OpenBLAS 0.3) are absent from the reference tree; they are restated from
 * the published netlib LAPACK/BLAS reference algorithms (dgbtf2, dpbtf2,
 * dpttrf, dgetf2, dgbtrs, dpbtrs/dtbsv, dptts2, dgetrs/dtrsm) in the same
 * operation order.
 *
 * Pinning (see tests/test_oracle_*.py): LAPACK factorisations and solves vs
 * scipy.linalg.lapack on the reference's own solver test matrices; basis
 * values vs scipy BSpline / closed forms / partition of unity; builds vs
 * scipy make_interp_spline and polynomial reproduction; integrals vs closed
 * forms; adjoints vs unit-vector Jacobians.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define OR_PERIODIC 0
#define OR_HERMITE 1
#define OR_HOMOGENEOUS_HERMITE 2
#define OR_GREVILLE 3

#define OR_MAXDEG 9

static char g_err[512];
const char* or_last_error(void) { return g_err; }
static void set_err(const char* m) { snprintf(g_err, sizeof g_err, "%s", m); }

/* ------------------------------------------------------------------------- */
/* B-splines                                                                   */
/* ------------------------------------------------------------------------- */

typedef struct {
    int degree, uniform, periodic, ncells;
    double origin, step; /* uniform knots: coordinate(k) = origin + k*step */
    double* knots;       /* non-uniform knots (ncells + 1 + 2*degree)       */
} or_bspl;

/* Uniform knots: UniformPointSampling::init_ghosted
 * (uniform_point_sampling.hpp:163-192) with n = ncells+1 and degree ghosts
 * each side (bsplines_uniform.hpp:135-148); coordinate() at :124-128. */
static double knot(const or_bspl* b, int k)
{
    if (b->uniform) return b->origin + (double)k * b->step;
    return b->knots[k];
}
static double bs_rmin(const or_bspl* b) { return knot(b, b->degree); }
static double bs_rmax(const or_bspl* b) { return knot(b, b->degree + b->ncells); }
static double bs_length(const or_bspl* b) { return bs_rmax(b) - bs_rmin(b); }
static int bs_nbasis(const or_bspl* b) { return b->ncells + (b->periodic ? 0 : b->degree); }
static int bs_size(const or_bspl* b) { return b->ncells + b->degree; }

static int bs_init_uniform(or_bspl* b, int degree, int periodic, double rmin, double rmax, int ncells)
{
    memset(b, 0, sizeof *b);
    b->degree = degree;
    b->uniform = 1;
    b->periodic = periodic;
    b->ncells = ncells;
    /* init_ghosted: step = (b - a) / (n - 1), origin = a - n_ghosts_before * step */
    double const step = (rmax - rmin) / (double)ncells;
    b->step = step;
    b->origin = rmin - (double)degree * step;
    return 0;
}

/* NonUniformBSplines::Impl ctor, bsplines_non_uniform.hpp:414-457 */
static int bs_init_nonuniform(or_bspl* b, int degree, int periodic, const double* breaks, int nbreaks)
{
    memset(b, 0, sizeof *b);
    b->degree = degree;
    b->uniform = 0;
    b->periodic = periodic;
    b->ncells = nbreaks - 1;
    int const nk = nbreaks + 2 * degree;
    b->knots = (double*)malloc(sizeof(double) * nk);
    for (int i = 0; i < nbreaks; ++i) {
        if (i > 0 && !(breaks[i] > breaks[i - 1])) { set_err("breaks not sorted"); return -1; }
        b->knots[degree + i] = breaks[i];
    }
    double const rmin = b->knots[degree];
    double const rmax = b->knots[nbreaks + degree - 1];
    if (periodic) {
        double const period = rmax - rmin;
        for (int i = 1; i < degree + 1; ++i) {
            b->knots[degree - i] = b->knots[degree + b->ncells - i] - period;
            b->knots[degree + b->ncells + i] = b->knots[degree + i] + period;
        }
    } else {
        for (int i = 1; i < degree + 1; ++i) {
            b->knots[degree - i] = rmin;
            b->knots[degree + nbreaks - 1 + i] = rmax;
        }
    }
    return 0;
}

/* UniformBSplines::get_icell_and_offset, bsplines_uniform.hpp:568-595 */
static void u_icell_offset(const or_bspl* b, double x, int* icell, double* offset)
{
    double const inv_dx = 1.0 / b->step;
    if (x <= bs_rmin(b)) {
        *icell = 0;
        *offset = 0.0;
    } else if (x >= bs_rmax(b)) {
        *icell = b->ncells - 1;
        *offset = 1.0;
    } else {
        double off = (x - bs_rmin(b)) * inv_dx;
        int ic = (int)off;
        off = off - ic;
        if (ic == b->ncells && off == 0.0) {
            ic = b->ncells - 1;
            off = 1.0;
        }
        *icell = ic;
        *offset = off;
    }
}

/* UniformBSplines eval_basis, bsplines_uniform.hpp:394-426 (nvals = degree+1
 * normally; integrals() calls it with degree+2 values). */
static int u_eval_basis(const or_bspl* b, double x, double* values, int nvals)
{
    double offset;
    int jmin;
    u_icell_offset(b, x, &jmin, &offset);
    values[0] = 1.0;
    for (int j = 1; j < nvals; ++j) {
        double xx = -offset;
        double saved = 0.0;
        for (int r = 0; r < j; ++r) {
            xx += 1;
            double const temp = values[r] / j;
            values[r] = saved + xx * temp;
            saved = (j - xx) * temp;
        }
        values[j] = saved;
    }
    return jmin;
}

/* UniformBSplines eval_basis_and_n_derivs, bsplines_uniform.hpp:475-564.
 * derivs is (degree+1) x (n+1), row-major: derivs[i*(n+1)+k]. */
static int u_eval_basis_and_n_derivs(const or_bspl* b, double x, int n, double* derivs)
{
    int const D = b->degree;
    double ndu[OR_MAXDEG + 1][OR_MAXDEG + 1];
    double a[OR_MAXDEG + 1][2];
    double offset;
    int jmin;
    u_icell_offset(b, x, &jmin, &offset);
    ndu[0][0] = 1.0;
    for (int j = 1; j < D + 1; ++j) {
        double xx = -offset;
        double saved = 0.0;
        for (int r = 0; r < j; ++r) {
            xx += 1.0;
            double const temp = ndu[j - 1][r] / j;
            ndu[j][r] = saved + xx * temp;
            saved = (j - xx) * temp;
        }
        ndu[j][j] = saved;
    }
    int const ld = n + 1;
    for (int i = 0; i < D + 1; ++i) derivs[i * ld + 0] = ndu[D][i];
    for (int r = 0; r < D + 1; ++r) {
        int s1 = 0, s2 = 1;
        a[0][0] = 1.0;
        for (int k = 1; k < n + 1; ++k) {
            double d = 0.0;
            int const rk = r - k;
            int const pk = D - k;
            if (r >= k) {
                a[0][s2] = a[0][s1] / (pk + 1);
                d = a[0][s2] * ndu[pk][rk];
            }
            int const j1 = rk > -1 ? 1 : (-rk);
            int const j2 = (r - 1) <= pk ? k : (D - r + 1);
            for (int j = j1; j < j2; ++j) {
                a[j][s2] = (a[j][s1] - a[j - 1][s1]) / (pk + 1);
                d += a[j][s2] * ndu[pk][rk + j];
            }
            if (r <= pk) {
                a[k][s2] = -a[k - 1][s1] / (pk + 1);
                d += a[k][s2] * ndu[pk][r];
            }
            derivs[r * ld + k] = d;
            int const t = s1; s1 = s2; s2 = t;
        }
    }
    double const inv_dx = 1.0 / b->step;
    double d = D * inv_dx;
    for (int k = 1; k < n + 1; ++k) {
        for (int i = 0; i < D + 1; ++i) derivs[i * ld + k] *= d;
        d *= (D - k) * inv_dx;
    }
    return jmin;
}

/* NonUniformBSplines find_cell_start, bsplines_non_uniform.hpp:668-695.
 * Returns the knot index of the cell start. */
static int nu_find_cell(const or_bspl* b, double x)
{
    int const front = b->degree, back = b->degree + b->ncells;
    if (x <= b->knots[front]) return front;
    if (x >= b->knots[back]) return back - 1;
    int low = front, high = back;
    int icell = low + (high - low) / 2;
    while (x < b->knots[icell] || x >= b->knots[icell + 1]) {
        if (x < b->knots[icell]) high = icell;
        else low = icell;
        icell = low + (high - low) / 2;
    }
    return icell;
}

/* NonUniformBSplines eval_basis, bsplines_non_uniform.hpp:461-497 */
static int nu_eval_basis(const or_bspl* b, double x, double* values)
{
    int const D = b->degree;
    double left[OR_MAXDEG], right[OR_MAXDEG];
    int const icell = nu_find_cell(b, x);
    values[0] = 1.0;
    for (int j = 0; j < D; ++j) {
        left[j] = x - b->knots[icell - j];
        right[j] = b->knots[icell + j + 1] - x;
        double saved = 0.0;
        for (int r = 0; r < j + 1; ++r) {
            double const temp = values[r] / (right[r] + left[j - r]);
            values[r] = saved + right[r] * temp;
            saved = left[j - r] * temp;
        }
        values[j + 1] = saved;
    }
    return icell - D;
}

/* NonUniformBSplines eval_basis_and_n_derivs, bsplines_non_uniform.hpp:561-664 */
static int nu_eval_basis_and_n_derivs(const or_bspl* b, double x, int n, double* derivs)
{
    int const D = b->degree;
    double left[OR_MAXDEG], right[OR_MAXDEG];
    double a[OR_MAXDEG + 1][2];
    double ndu[OR_MAXDEG + 1][OR_MAXDEG + 1];
    int const icell = nu_find_cell(b, x);
    ndu[0][0] = 1.0;
    for (int j = 0; j < D; ++j) {
        left[j] = x - b->knots[icell - j];
        right[j] = b->knots[icell + j + 1] - x;
        double saved = 0.0;
        for (int r = 0; r < j + 1; ++r) {
            ndu[r][j + 1] = 1.0 / (right[r] + left[j - r]);
            double const temp = ndu[j][r] * ndu[r][j + 1];
            ndu[j + 1][r] = saved + right[r] * temp;
            saved = left[j - r] * temp;
        }
        ndu[j + 1][j + 1] = saved;
    }
    int const ld = n + 1;
    for (int j = 0; j < D + 1; ++j) derivs[j * ld + 0] = ndu[D][j];
    for (int r = 0; r < D + 1; ++r) {
        int s1 = 0, s2 = 1;
        a[0][0] = 1.0;
        for (int k = 1; k < n + 1; ++k) {
            double d = 0.0;
            int const rk = r - k;
            int const pk = D - k;
            if (r >= k) {
                a[0][s2] = a[0][s1] * ndu[rk][pk + 1];
                d = a[0][s2] * ndu[pk][rk];
            }
            int const j1 = rk > -1 ? 1 : (-rk);
            int const j2 = (r - 1) <= pk ? k : (D - r + 1);
            for (int j = j1; j < j2; ++j) {
                a[j][s2] = (a[j][s1] - a[j - 1][s1]) * ndu[rk + j][pk + 1];
                d += a[j][s2] * ndu[pk][rk + j];
            }
            if (r <= pk) {
                a[k][s2] = -a[k - 1][s1] * ndu[r][pk + 1];
                d += a[k][s2] * ndu[pk][r];
            }
            derivs[r * ld + k] = d;
            int const t = s1; s1 = s2; s2 = t;
        }
    }
    int r = D;
    for (int k = 1; k < n + 1; ++k) {
        for (int i = 0; i < D + 1; ++i) derivs[i * ld + k] *= r;
        r *= D - k;
    }
    return icell - D;
}

static int bs_eval_basis(const or_bspl* b, double x, double* values)
{
    return b->uniform ? u_eval_basis(b, x, values, b->degree + 1) : nu_eval_basis(b, x, values);
}
static int bs_eval_basis_and_n_derivs(const or_bspl* b, double x, int n, double* derivs)
{
    return b->uniform ? u_eval_basis_and_n_derivs(b, x, n, derivs)
                      : nu_eval_basis_and_n_derivs(b, x, n, derivs);
}

/* integrals(), integrals.hpp:30-101 (uniform) and :110-153 (non-uniform);
 * fills the full spline domain (size() entries, zeros on periodic repeats). */
static void bs_integrals(const or_bspl* b, double* out)
{
    int const D = b->degree, nb = bs_nbasis(b), sz = bs_size(b);
    for (int i = 0; i < sz; ++i) out[i] = 0.0;
    if (b->uniform) {
        if (b->periodic) {
            for (int i = 0; i < nb; ++i) out[i] = b->step;
        } else {
            for (int i = D; i < sz - D; ++i) out[i] = b->step;
            double ev[OR_MAXDEG + 2];
            for (int i = 0; i < D; ++i) {
                u_eval_basis(b, bs_rmin(b), ev, D + 2);
                double d_eval = 0.0;
                for (int k = 0; k < D + 2; ++k) d_eval += ev[k];
                double c_eval = 0.0;
                for (int k = 0; k < D - i; ++k) c_eval += ev[k];
                double const edge = b->step * (d_eval - c_eval);
                out[i] = edge;
                out[sz - 1 - i] = edge;
            }
        }
    } else {
        double const inv_deg = 1.0 / (D + 1);
        for (int i = 0; i < nb; ++i) out[i] = (b->knots[i + D + 1] - b->knots[i]) * inv_deg;
    }
}

/* ------------------------------------------------------------------------- */
/* Linear problems (netlib LAPACK/BLAS algorithm order)                        */
/* ------------------------------------------------------------------------- */

enum { P_DENSE, P_BAND, P_PDSBAND, P_PDSTRIDIAG, P_2X2, P_3X3 };

typedef struct or_prob {
    int type, n;
    /* dense: a (n x n row-major), ipiv */
    double* a;
    int* ipiv;
    /* band: q ((2kl+ku+1) x n row-major); pds band: q ((kd+1) x n); tridiag: q (2 x n) */
    int kl, ku;
    double* q;
    /* blocks */
    struct or_prob *tl, *br;
    double *tr, *bl; /* tr: nq x nb, bl: nb x nq (row-major) */
    int tr_nnz, bl_nnz;
    int *tr_r, *tr_c, *bl_r, *bl_c;
    double *tr_v, *bl_v;
    int top;
} or_prob;

static or_prob* prob_new(int type, int n)
{
    or_prob* p = (or_prob*)calloc(1, sizeof(or_prob));
    p->type = type;
    p->n = n;
    return p;
}

static void prob_free(or_prob* p)
{
    if (!p) return;
    free(p->a); free(p->ipiv); free(p->q);
    prob_free(p->tl); prob_free(p->br);
    free(p->tr); free(p->bl);
    free(p->tr_r); free(p->tr_c); free(p->tr_v);
    free(p->bl_r); free(p->bl_c); free(p->bl_v);
    free(p);
}

/* splines_linear_problem_dense.cpp:35-41 */
static or_prob* make_dense(int n)
{
    or_prob* p = prob_new(P_DENSE, n);
    p->a = (double*)calloc((size_t)n * n + 1, sizeof(double));
    p->ipiv = (int*)calloc(n + 1, sizeof(int));
    return p;
}
/* splines_linear_problem_band.cpp:31-49 */
static or_prob* make_band(int n, int kl, int ku)
{
    or_prob* p = prob_new(P_BAND, n);
    p->kl = kl; p->ku = ku;
    p->q = (double*)calloc((size_t)(2 * kl + ku + 1) * n, sizeof(double));
    p->ipiv = (int*)calloc(n + 1, sizeof(int));
    return p;
}
/* splines_linear_problem_pds_band.cpp:31-40 */
static or_prob* make_pdsband(int n, int kd)
{
    or_prob* p = prob_new(P_PDSBAND, n);
    p->kl = kd; p->ku = kd;
    p->q = (double*)calloc((size_t)(kd + 1) * n, sizeof(double));
    return p;
}
/* splines_linear_problem_pds_tridiag.cpp:31-37 */
static or_prob* make_pdstridiag(int n)
{
    or_prob* p = prob_new(P_PDSTRIDIAG, n);
    p->kl = 1; p->ku = 1;
    p->q = (double*)calloc((size_t)2 * n, sizeof(double));
    return p;
}
/* splines_linear_problem_2x2_blocks.cpp:76-96 */
static or_prob* make_2x2(int n, or_prob* tl)
{
    or_prob* p = prob_new(P_2X2, n);
    p->tl = tl;
    int const nq = tl->n, nb = n - nq;
    p->tr = (double*)calloc((size_t)nq * nb + 1, sizeof(double));
    p->bl = (double*)calloc((size_t)nb * nq + 1, sizeof(double));
    p->br = make_dense(nb);
    return p;
}
/* splines_linear_problem_3x3_blocks.cpp:18-26 */
static or_prob* make_3x3(int n, int top, or_prob* center)
{
    or_prob* p = make_2x2(n, center);
    p->type = P_3X3;
    p->top = top;
    return p;
}

/* splines_linear_problem_maker.hpp:56-76 */
static or_prob* maker_band(int n, int kl, int ku, int pds)
{
    if (kl == ku && kl == 1 && pds) return make_pdstridiag(n);
    if (kl == ku && pds) return make_pdsband(n, kl);
    if (2 * kl + ku + 1 >= n) return make_dense(n);
    return make_band(n, kl, ku);
}
/* splines_linear_problem_maker.hpp:92-111 */
static or_prob* maker_block_band(int n, int kl, int ku, int pds, int bottom, int top)
{
    int const main = n - top - bottom;
    or_prob* m = maker_band(main, kl, ku, pds);
    if (top == 0) return make_2x2(n, m);
    return make_3x3(n, top, m);
}
/* splines_linear_problem_maker.hpp:129-146 */
static or_prob* maker_periodic_band(int n, int kl, int ku, int pds)
{
    int const bottom = kl > ku ? kl : ku;
    int const top = n - bottom;
    if ((long)bottom * (n + top) + (long)(2 * kl + ku + 1) * top >= (long)n * n) return make_dense(n);
    return maker_block_band(n, kl, ku, pds, bottom, 0);
}

static double prob_get(const or_prob* p, int i, int j);

/* 3x3 index remap, splines_linear_problem_3x3_blocks.cpp:31-47 */
static void adjust3(const or_prob* p, int* i, int* j)
{
    int const nq = p->tl->n;
    if (*i < p->top) *i += nq;
    else if (*i < p->top + nq) *i -= p->top;
    if (*j < p->top) *j += nq;
    else if (*j < p->top + nq) *j -= p->top;
}

static void prob_set(or_prob* p, int i, int j, double v)
{
    switch (p->type) {
    case P_DENSE: p->a[(size_t)i * p->n + j] = v; break;
    case P_BAND: { /* band.cpp:85-106 */
        int const kl = p->kl, ku = p->ku;
        if (i >= (j - ku > 0 ? j - ku : 0) && i < (p->n < j + kl + 1 ? p->n : j + kl + 1))
            p->q[(size_t)(kl + ku + i - j) * p->n + j] = v;
        break;
    }
    case P_PDSBAND: { /* pds_band.cpp:64-81 */
        if (i > j) { int t = i; i = j; j = t; }
        if (j - i < p->kl + 1) p->q[(size_t)(j - i) * p->n + i] = v;
        break;
    }
    case P_PDSTRIDIAG: { /* pds_tridiag.cpp:61-78 */
        if (i > j) { int t = i; i = j; j = t; }
        if (j - i < 2) p->q[(size_t)(j - i) * p->n + i] = v;
        break;
    }
    case P_3X3: adjust3(p, &i, &j); /* fallthrough */
    case P_2X2: { /* 2x2_blocks.cpp:127-145 */
        int const nq = p->tl->n, nb = p->n - nq;
        if (i < nq && j < nq) prob_set(p->tl, i, j, v);
        else if (i >= nq && j >= nq) prob_set(p->br, i - nq, j - nq, v);
        else if (j >= nq) p->tr[(size_t)i * nb + (j - nq)] = v;
        else p->bl[(size_t)(i - nq) * nq + j] = v;
        break;
    }
    }
}

static double prob_get(const or_prob* p, int i, int j)
{
    switch (p->type) {
    case P_DENSE: return p->a[(size_t)i * p->n + j];
    case P_BAND: {
        int const kl = p->kl, ku = p->ku;
        if (i >= (j - ku > 0 ? j - ku : 0) && i < (p->n < j + kl + 1 ? p->n : j + kl + 1))
            return p->q[(size_t)(kl + ku + i - j) * p->n + j];
        return 0.0;
    }
    case P_PDSBAND:
        if (i > j) { int t = i; i = j; j = t; }
        if (j - i < p->kl + 1) return p->q[(size_t)(j - i) * p->n + i];
        return 0.0;
    case P_PDSTRIDIAG:
        if (i > j) { int t = i; i = j; j = t; }
        if (j - i < 2) return p->q[(size_t)(j - i) * p->n + i];
        return 0.0;
    case P_3X3: adjust3(p, &i, &j); /* fallthrough */
    case P_2X2: {
        int const nq = p->tl->n, nb = p->n - nq;
        if (i < nq && j < nq) return prob_get(p->tl, i, j);
        if (i >= nq && j >= nq) return prob_get(p->br, i - nq, j - nq);
        if (j >= nq) return p->tr[(size_t)i * nb + (j - nq)];
        return p->bl[(size_t)(i - nq) * nq + j];
    }
    }
    return 0.0;
}

/* ---- per-column solves: b is one right-hand side with element stride ld ---- */

/* dptts2 (called by dpttrs), LDL^T of a tridiagonal SPD matrix */
static void pttrs_col(const or_prob* p, double* b, long ld)
{
    int const n = p->n;
    const double* d = p->q;
    const double* e = p->q + n;
    if (n == 1) { b[0] = b[0] * (1.0 / d[0]); return; }
    for (int i = 1; i < n; ++i) b[i * ld] = b[i * ld] - b[(i - 1) * ld] * e[i - 1];
    b[(n - 1) * ld] = b[(n - 1) * ld] / d[n - 1];
    for (int i = n - 2; i >= 0; --i) b[i * ld] = b[i * ld] / d[i] - b[(i + 1) * ld] * e[i];
}

/* dpbtrs('L') = dtbsv('L','N','N') then dtbsv('L','T','N'); band AB(r, c) = q[r*n + c] */
static void pbtrs_col(const or_prob* p, double* x, long ld)
{
    int const n = p->n, k = p->kl;
    const double* q = p->q;
    for (int j = 0; j < n; ++j) {
        if (x[j * ld] != 0.0) {
            x[j * ld] = x[j * ld] / q[j];
            double const temp = x[j * ld];
            int const iend = n - 1 < j + k ? n - 1 : j + k;
            for (int i = j + 1; i <= iend; ++i) x[i * ld] = x[i * ld] - temp * q[(size_t)(i - j) * n + j];
        }
    }
    for (int j = n - 1; j >= 0; --j) {
        double temp = x[j * ld];
        int const iend = n - 1 < j + k ? n - 1 : j + k;
        for (int i = iend; i >= j + 1; --i) temp = temp - q[(size_t)(i - j) * n + j] * x[i * ld];
        temp = temp / q[j];
        x[j * ld] = temp;
    }
}

/* dgbtrs; AB(r, c) = q[r*n + c], kv = kl+ku */
static void gbtrs_col(const or_prob* p, double* b, long ld, int trans)
{
    int const n = p->n, kl = p->kl, ku = p->ku, kd = kl + ku;
    const double* q = p->q;
#define AB(r, c) q[(size_t)(r) * n + (c)]
    if (!trans) {
        if (kl > 0) {
            for (int j = 0; j < n - 1; ++j) {
                int const lm = kl < n - 1 - j ? kl : n - 1 - j;
                int const l = p->ipiv[j];
                if (l != j) { double t = b[l * ld]; b[l * ld] = b[j * ld]; b[j * ld] = t; }
                double const temp = -1.0 * b[j * ld];
                if (b[j * ld] != 0.0)
                    for (int i = 0; i < lm; ++i) b[(j + 1 + i) * ld] = b[(j + 1 + i) * ld] + AB(kd + 1 + i, j) * temp;
            }
        }
        /* dtbsv('U','N','N', n, kl+ku) */
        for (int j = n - 1; j >= 0; --j) {
            if (b[j * ld] != 0.0) {
                b[j * ld] = b[j * ld] / AB(kd, j);
                double const temp = b[j * ld];
                int const i0 = j - kd > 0 ? j - kd : 0;
                for (int i = j - 1; i >= i0; --i) b[i * ld] = b[i * ld] - temp * AB(kd + i - j, j);
            }
        }
    } else {
        /* dtbsv('U','T','N', n, kl+ku) */
        for (int j = 0; j < n; ++j) {
            double temp = b[j * ld];
            int const i0 = j - kd > 0 ? j - kd : 0;
            for (int i = i0; i <= j - 1; ++i) temp = temp - AB(kd + i - j, j) * b[i * ld];
            temp = temp / AB(kd, j);
            b[j * ld] = temp;
        }
        if (kl > 0) {
            for (int j = n - 2; j >= 0; --j) {
                int const lm = kl < n - 1 - j ? kl : n - 1 - j;
                double temp = 0.0;
                for (int i = 0; i < lm; ++i) temp = temp + AB(kd + 1 + i, j) * b[(j + 1 + i) * ld];
                b[j * ld] = b[j * ld] + -1.0 * temp;
                int const l = p->ipiv[j];
                if (l != j) { double t = b[l * ld]; b[l * ld] = b[j * ld]; b[j * ld] = t; }
            }
        }
    }
#undef AB
}

/* dgetrs (dlaswp + dtrsm L unit + dtrsm U) on row-major a */
static void getrs_col(const or_prob* p, double* b, long ld, int trans)
{
    int const n = p->n;
    const double* a = p->a;
    if (n == 0) return; /* dense.cpp:98-100 */
#define A(i, j) a[(size_t)(i) * n + (j)]
    if (!trans) {
        for (int j = 0; j < n; ++j) {
            int const l = p->ipiv[j];
            if (l != j) { double t = b[l * ld]; b[l * ld] = b[j * ld]; b[j * ld] = t; }
        }
        for (int j = 0; j < n; ++j)
            if (b[j * ld] != 0.0)
                for (int i = j + 1; i < n; ++i) b[i * ld] = b[i * ld] - b[j * ld] * A(i, j);
        for (int j = n - 1; j >= 0; --j) {
            if (b[j * ld] != 0.0) {
                b[j * ld] = b[j * ld] / A(j, j);
                for (int i = 0; i < j; ++i) b[i * ld] = b[i * ld] - b[j * ld] * A(i, j);
            }
        }
    } else {
        for (int j = 0; j < n; ++j) {
            double temp = b[j * ld];
            for (int i = 0; i < j; ++i) temp = temp - A(i, j) * b[i * ld];
            temp = temp / A(j, j);
            b[j * ld] = temp;
        }
        for (int j = n - 1; j >= 0; --j) {
            double temp = b[j * ld];
            for (int i = j + 1; i < n; ++i) temp = temp - A(i, j) * b[i * ld];
            b[j * ld] = temp;
        }
        for (int j = n - 1; j >= 0; --j) {
            int const l = p->ipiv[j];
            if (l != j) { double t = b[l * ld]; b[l * ld] = b[j * ld]; b[j * ld] = t; }
        }
    }
#undef A
}

static void prob_solve_col(const or_prob* p, double* b, long ld, int trans);

/* 2x2_blocks.cpp:237-298 for one column */
static void solve2x2_col(const or_prob* p, double* b, long ld, int trans)
{
    int const nq = p->tl->n;
    double* b1 = b;
    double* b2 = b + (long)nq * ld;
    if (!trans) {
        prob_solve_col(p->tl, b1, ld, 0);
        for (int z = 0; z < p->bl_nnz; ++z) b2[p->bl_r[z] * ld] -= p->bl_v[z] * b1[p->bl_c[z] * ld];
        prob_solve_col(p->br, b2, ld, 0);
        for (int z = 0; z < p->tr_nnz; ++z) b1[p->tr_r[z] * ld] -= p->tr_v[z] * b2[p->tr_c[z] * ld];
    } else {
        for (int z = 0; z < p->tr_nnz; ++z) b2[p->tr_c[z] * ld] -= p->tr_v[z] * b1[p->tr_r[z] * ld];
        prob_solve_col(p->br, b2, ld, 1);
        for (int z = 0; z < p->bl_nnz; ++z) b1[p->bl_c[z] * ld] -= p->bl_v[z] * b2[p->bl_r[z] * ld];
        prob_solve_col(p->tl, b1, ld, 1);
    }
}

/* 3x3_blocks.cpp:66-145 for one column: b has n + top rows */
static void solve3x3_col(const or_prob* p, double* b, long ld, int trans)
{
    int const nq = p->tl->n, top = p->top, n = p->n;
    int const nbot = n - top - nq;
    double tmp[64];
    /* interchange_rows_from_3_to_2_blocks_rhs */
    for (int i = 0; i < nbot; ++i) tmp[i] = b[(long)(top + nq + i) * ld];
    for (int i = 0; i < nbot; ++i) b[(long)(2 * top + nq + i) * ld] = tmp[i];
    for (int i = 0; i < top; ++i) b[(long)(top + nq + i) * ld] = b[(long)i * ld];
    solve2x2_col(p, b + (long)top * ld, ld, trans);
    /* interchange_rows_from_2_to_3_blocks_rhs */
    for (int i = 0; i < top; ++i) b[(long)i * ld] = b[(long)(top + nq + i) * ld];
    for (int i = 0; i < nbot; ++i) tmp[i] = b[(long)(2 * top + nq + i) * ld];
    for (int i = 0; i < nbot; ++i) b[(long)(top + nq + i) * ld] = tmp[i];
}

static void prob_solve_col(const or_prob* p, double* b, long ld, int trans)
{
    switch (p->type) {
    case P_DENSE: getrs_col(p, b, ld, trans); break;
    case P_BAND: gbtrs_col(p, b, ld, trans); break;
    case P_PDSBAND: pbtrs_col(p, b, ld); break;
    case P_PDSTRIDIAG: pttrs_col(p, b, ld); break;
    case P_2X2: solve2x2_col(p, b, ld, trans); break;
    case P_3X3: solve3x3_col(p, b, ld, trans); break;
    }
}

static int prob_rhs_rows(const or_prob* p) { return p->type == P_3X3 ? p->n + p->top : p->n; }

/* ---- factorisations ---- */

/* dpttrf */
static int pttrf(or_prob* p)
{
    int const n = p->n;
    double* d = p->q;
    double* e = p->q + n;
    for (int i = 0; i < n - 1; ++i) {
        if (d[i] <= 0.0) return i + 1;
        double const ei = e[i];
        e[i] = ei / d[i];
        d[i + 1] = d[i + 1] - e[i] * ei;
    }
    if (n > 0 && d[n - 1] <= 0.0) return n;
    return 0;
}

/* dpbtrf('L') -> dpbtf2 (ILAENV gives NB=1 for KD<=64) */
static int pbtrf(or_prob* p)
{
    int const n = p->n, kd = p->kl;
    double* q = p->q;
#define AB(r, c) q[(size_t)(r) * n + (c)]
    for (int j = 0; j < n; ++j) {
        double ajj = AB(0, j);
        if (ajj <= 0.0) return j + 1;
        ajj = sqrt(ajj);
        AB(0, j) = ajj;
        int const kn = kd < n - 1 - j ? kd : n - 1 - j;
        if (kn > 0) {
            double const s = 1.0 / ajj;
            for (int i = 1; i <= kn; ++i) AB(i, j) = s * AB(i, j);
            /* dsyr('L', kn, -1, x = AB(1..kn, j), A = trailing block) */
            for (int jj = 0; jj < kn; ++jj) {
                double const xj = AB(1 + jj, j);
                if (xj != 0.0) {
                    double const temp = -1.0 * xj;
                    for (int ii = jj; ii < kn; ++ii)
                        AB(ii - jj, j + 1 + jj) = AB(ii - jj, j + 1 + jj) + AB(1 + ii, j) * temp;
                }
            }
        }
    }
#undef AB
    return 0;
}

/* dgbtrf -> dgbtf2 (ILAENV gives NB=1 for KU<=64); ipiv 0-based */
static int gbtrf(or_prob* p)
{
    int const n = p->n, kl = p->kl, ku = p->ku, kv = ku + kl;
    double* q = p->q;
    int info = 0;
#define AB(r, c) q[(size_t)(r) * n + (c)]
    for (int j = ku + 1; j < (kv < n ? kv : n); ++j)
        for (int i = kv - j; i < kl; ++i) AB(i, j) = 0.0;
    int ju = 0;
    for (int j = 0; j < n; ++j) {
        if (j + kv < n)
            for (int i = 0; i < kl; ++i) AB(i, j + kv) = 0.0;
        int const km = kl < n - 1 - j ? kl : n - 1 - j;
        int jp = 0;
        double amax = fabs(AB(kv, j));
        for (int i = 1; i <= km; ++i)
            if (fabs(AB(kv + i, j)) > amax) { amax = fabs(AB(kv + i, j)); jp = i; }
        p->ipiv[j] = jp + j;
        if (AB(kv + jp, j) != 0.0) {
            int const t = j + ku + jp < n - 1 ? j + ku + jp : n - 1;
            ju = ju > t ? ju : t;
            if (jp != 0)
                for (int c = j; c <= ju; ++c) {
                    double const tmp = AB(kv + jp + j - c, c);
                    AB(kv + jp + j - c, c) = AB(kv + j - c, c);
                    AB(kv + j - c, c) = tmp;
                }
            if (km > 0) {
                double const s = 1.0 / AB(kv, j);
                for (int i = 1; i <= km; ++i) AB(kv + i, j) = s * AB(kv + i, j);
                for (int c = j + 1; c <= ju; ++c) {
                    double const y = AB(kv + j - c, c);
                    if (y != 0.0) {
                        double const temp = -1.0 * y;
                        for (int i = 1; i <= km; ++i) AB(kv + j + i - c, c) = AB(kv + j + i - c, c) + AB(kv + i, j) * temp;
                    }
                }
            }
        } else if (info == 0) {
            info = j + 1;
        }
    }
#undef AB
    return info;
}

/* dgetrf (unblocked right-looking dgetf2 order); ipiv 0-based */
static int getrf(or_prob* p)
{
    int const n = p->n;
    double* a = p->a;
    int info = 0;
#define A(i, j) a[(size_t)(i) * n + (j)]
    for (int j = 0; j < n; ++j) {
        int jp = j;
        double amax = fabs(A(j, j));
        for (int i = j + 1; i < n; ++i)
            if (fabs(A(i, j)) > amax) { amax = fabs(A(i, j)); jp = i; }
        p->ipiv[j] = jp;
        if (A(jp, j) != 0.0) {
            if (jp != j)
                for (int c = 0; c < n; ++c) { double t = A(j, c); A(j, c) = A(jp, c); A(jp, c) = t; }
            if (j < n - 1) {
                double const s = 1.0 / A(j, j);
                for (int i = j + 1; i < n; ++i) A(i, j) = s * A(i, j);
            }
        } else if (info == 0) {
            info = j + 1;
        }
        for (int c = j + 1; c < n; ++c) {
            double const y = A(j, c);
            if (y != 0.0) {
                double const temp = -1.0 * y;
                for (int i = j + 1; i < n; ++i) A(i, c) = A(i, c) + A(i, j) * temp;
            }
        }
    }
#undef A
    return info;
}

static int prob_setup(or_prob* p);

/* dense2coo, 2x2_blocks.cpp:147-190 (row-major scan, |a| >= tol) */
static int dense2coo(const double* m, int nr, int nc, double tol, int** r, int** c, double** v)
{
    int nnz = 0;
    *r = (int*)malloc(sizeof(int) * ((size_t)nr * nc + 1));
    *c = (int*)malloc(sizeof(int) * ((size_t)nr * nc + 1));
    *v = (double*)malloc(sizeof(double) * ((size_t)nr * nc + 1));
    for (int i = 0; i < nr; ++i)
        for (int j = 0; j < nc; ++j) {
            double const aij = m[(size_t)i * nc + j];
            if (fabs(aij) >= tol) {
                (*r)[nnz] = i;
                (*c)[nnz] = j;
                (*v)[nnz] = aij;
                ++nnz;
            }
        }
    return nnz;
}

/* 2x2_blocks.cpp:193-234 */
static int setup2x2(or_prob* p)
{
    int info = prob_setup(p->tl);
    if (info) return info;
    int const nq = p->tl->n, nb = p->n - nq;
    for (int j = 0; j < nb; ++j) prob_solve_col(p->tl, p->tr + j, nb, 0);
    p->tr_nnz = dense2coo(p->tr, nq, nb, 1e-14, &p->tr_r, &p->tr_c, &p->tr_v);
    p->bl_nnz = dense2coo(p->bl, nb, nq, 1e-14, &p->bl_r, &p->bl_c, &p->bl_v);
    for (int i = 0; i < nb; ++i)
        for (int j = 0; j < nb; ++j) {
            double val = 0.0;
            for (int l = 0; l < nq; ++l) val += p->bl[(size_t)i * nq + l] * p->tr[(size_t)l * nb + j];
            prob_set(p->br, i, j, prob_get(p->br, i, j) - val);
        }
    return prob_setup(p->br);
}

static int prob_setup(or_prob* p)
{
    switch (p->type) {
    case P_DENSE: return getrf(p);
    case P_BAND: return gbtrf(p);
    case P_PDSBAND: return pbtrf(p);
    case P_PDSTRIDIAG: return pttrf(p);
    case P_2X2:
    case P_3X3: return setup2x2(p);
    }
    return -1;
}

/* ------------------------------------------------------------------------- */
/* Public oracle objects: problems (for solver known-answer tests)            */
/* ------------------------------------------------------------------------- */

void* or_problem_create(int kind, int n, int kl, int ku, int pds, int bottom, int top)
{
    /* kind: 0 dense, 1 maker_band, 2 block (2x2/3x3 with band main), 3 periodic band,
     *       4 2x2 with a Dense main block, 5 3x3 with a Dense center block,
     *       6 SplinesLinearProblemBand(n, kl, ku) built directly */
    switch (kind) {
    case 6: return make_band(n, kl, ku);
    case 0: return make_dense(n);
    case 1: return maker_band(n, kl, ku, pds);
    case 2: return maker_block_band(n, kl, ku, pds, bottom, top);
    case 3: return maker_periodic_band(n, kl, ku, pds);
    case 4: return make_2x2(n, make_dense(n - bottom));
    case 5: return make_3x3(n, top, make_dense(n - top - bottom));
    }
    return NULL;
}
void or_problem_set(void* h, int i, int j, double v) { prob_set((or_prob*)h, i, j, v); }
double or_problem_get(void* h, int i, int j) { return prob_get((or_prob*)h, i, j); }
int or_problem_setup(void* h) { return prob_setup((or_prob*)h); }
int or_problem_type(void* h) { return ((or_prob*)h)->type; }
int or_problem_rhs_rows(void* h) { return prob_rhs_rows((or_prob*)h); }
/* b: (rhs_rows x nrhs) LayoutRight, as SplinesLinearProblem::solve */
void or_problem_solve(void* h, double* b, int nrhs, int trans)
{
    or_prob* p = (or_prob*)h;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nrhs; ++j) prob_solve_col(p, b + j, nrhs, trans);
}
void or_problem_destroy(void* h) { prob_free((or_prob*)h); }
/* Raw factor storage access for LAPACK comparison tests. */
const double* or_problem_q(void* h) { or_prob* p = (or_prob*)h; return p->type == P_DENSE ? p->a : p->q; }
const int* or_problem_ipiv(void* h) { return ((or_prob*)h)->ipiv; }

/* ------------------------------------------------------------------------- */
/* Spline builder / evaluator                                                  */
/* ------------------------------------------------------------------------- */

typedef struct {
    or_bspl bs;
    int bcl, bcu;
    int nbe_l, nbe_r, nbv_l, nbv_r, odd;
    int npts;
    double* pts;
    double dx;
    int offset;
    or_prob* mat;
} or_spline;

static int n_boundary_equations(int bc, int degree)
{
    return (bc == OR_HERMITE || bc == OR_HOMOGENEOUS_HERMITE) ? degree / 2 : 0;
}

/* GrevilleInterpolationPoints, greville_interpolation_points.hpp:46-295 */
static int greville_points(const or_spline* s, double* out)
{
    const or_bspl* b = &s->bs;
    int const D = b->degree, nb = bs_nbasis(b);
    int const NBE_MIN = n_boundary_equations(s->bcl, D), NBE_MAX = n_boundary_equations(s->bcu, D);
    int const NBE = NBE_MIN + NBE_MAX;
    int const npoints = nb - NBE;
    int const uniform_dd = b->uniform && ((NBE_MIN != 0 && NBE_MAX != 0) || b->periodic);
    double const shift = (D % 2 == 0) ? 0.5 : 0.0;
    double const dxg = (bs_rmax(b) - bs_rmin(b)) / b->ncells;
    if (uniform_dd) {
        /* UniformPointSampling(origin = rmin + shift*dx, step = dx) */
        double const origin = bs_rmin(b) + shift * dxg;
        for (int i = 0; i < npoints; ++i) out[i] = origin + (double)i * dxg;
        return npoints;
    }
    if (b->uniform) {
        double const origin = bs_rmin(b) + shift * dxg;
        if (s->bcl == OR_GREVILLE) {
            for (int i = 0; i < D / 2 + 1; ++i) {
                out[i] = (D - i) * bs_rmin(b);
                /* knots [last_support_knot(i) - i, last_support_knot(i)) = [D+1, D+1+i) */
                for (int k = D + 1; k < D + 1 + i; ++k) out[i] += knot(b, k);
                out[i] /= D;
            }
        } else {
            out[0] = origin + 0.0 * dxg;
        }
        int const n_start = (s->bcl == OR_GREVILLE) ? D / 2 + 1 : 1;
        int const domain_size = b->ncells + 1 - 2;
        for (int ip = 1; ip < 1 + domain_size; ++ip) out[ip - 1 + n_start] = origin + (double)ip * dxg;
        if (s->bcu == OR_GREVILLE) {
            for (int i = 0; i < D / 2 + 1; ++i) {
                out[npoints - 1 - i] = (D - i) * bs_rmax(b);
                int const first = (nb - 1 - i) + 1;
                for (int k = first; k < first + i; ++k) out[npoints - 1 - i] += knot(b, k);
                out[npoints - 1 - i] /= D;
            }
        } else {
            out[npoints - 1] = origin + (double)(b->ncells - 1 + D % 2) * dxg;
        }
        return npoints;
    }
    /* non-uniform: averages of D knots */
    int const ng = b->periodic ? nb + 1 : nb;
    double* g = (double*)malloc(sizeof(double) * (ng + 1));
    for (int ib = 0; ib < nb; ++ib) {
        g[ib] = 0.0;
        for (int k = ib + 1; k < ib + 1 + D; ++k) g[ib] += knot(b, k);
        g[ib] /= D;
    }
    if (b->periodic) {
        double temp[OR_MAXDEG + 1];
        int np = 0;
        while (g[np] < bs_rmin(b)) {
            temp[np] = g[np] + bs_length(b);
            ++np;
        }
        for (int i = 0; i < nb - np; ++i) g[i] = g[i + np];
        for (int i = 0; i < np; ++i) g[nb - np + i] = temp[i];
        g[ng - 1] = g[0] + bs_length(b);
    }
    for (int i = 0; i < npoints; ++i) out[i] = g[NBE_MIN + i];
    free(g);
    return npoints;
}

/* check_valid_grid, spline_builder.hpp:1273-1317 */
static int check_valid_grid(const or_spline* s)
{
    const or_bspl* b = &s->bs;
    int const expected = bs_nbasis(b) - s->nbe_l - s->nbe_r;
    if (s->npts != expected) { set_err("Incorrect number of points supplied to NonUniformInterpolationPoints."); return -1; }
    int nin = 0;
    int cur = b->degree + 1; /* knot index of the current cell end */
    for (int i = 0; i < s->npts; ++i) {
        double const pt = s->pts[i];
        if (pt > knot(b, cur)) {
            if (nin > b->degree + 1) { set_err("The spline problem is overconstrained."); return -1; }
            nin = 1;
            cur += 1;
        } else if (pt == knot(b, cur)) {
            if (nin + 1 > b->degree + 1) { set_err("The spline problem is overconstrained."); return -1; }
            nin = 1;
            cur += 1;
        } else {
            nin += 1;
        }
    }
    if (nin > b->degree + 1) { set_err("The spline problem is overconstrained."); return -1; }
    return 0;
}

static int modulo_i(int x, int y) { return x - y * (int)floor((double)x / y); }

void or_spline_destroy(void* h)
{
    or_spline* s = (or_spline*)h;
    if (!s) return;
    free(s->bs.knots);
    free(s->pts);
    prob_free(s->mat);
    free(s);
}

/* SplineBuilder ctor, spline_builder.hpp:227-260, :627-886 */
void* or_spline_create(int degree, int uniform, int bcl, int bcu, int ncells, double rmin, double rmax,
                       const double* breaks, const double* pts, int npts_in)
{
    if (degree < 1 || degree > OR_MAXDEG) { set_err("degree out of range"); return NULL; }
    if ((bcl == OR_PERIODIC) != (bcu == OR_PERIODIC)) { set_err("Incompatible closure relations"); return NULL; }
    or_spline* s = (or_spline*)calloc(1, sizeof(or_spline));
    int const periodic = bcl == OR_PERIODIC;
    if (uniform) bs_init_uniform(&s->bs, degree, periodic, rmin, rmax, ncells);
    else if (bs_init_nonuniform(&s->bs, degree, periodic, breaks, ncells + 1)) { or_spline_destroy(s); return NULL; }
    const or_bspl* b = &s->bs;
    int const D = degree;
    s->bcl = bcl; s->bcu = bcu;
    s->nbe_l = n_boundary_equations(bcl, D);
    s->nbe_r = n_boundary_equations(bcu, D);
    s->nbv_l = bcl == OR_HOMOGENEOUS_HERMITE ? 0 : s->nbe_l;
    s->nbv_r = bcu == OR_HOMOGENEOUS_HERMITE ? 0 : s->nbe_r;
    s->odd = D % 2;
    int const nb = bs_nbasis(b);
    s->pts = (double*)malloc(sizeof(double) * (nb + 2));
    if (pts) {
        s->npts = npts_in;
        if (npts_in > nb) { set_err("too many interpolation points"); or_spline_destroy(s); return NULL; }
        memcpy(s->pts, pts, sizeof(double) * npts_in);
    } else {
        s->npts = greville_points(s, s->pts);
    }
    s->dx = (bs_rmax(b) - bs_rmin(b)) / b->ncells;
    if (check_valid_grid(s)) { or_spline_destroy(s); return NULL; }
    /* compute_offset, :627-658 */
    s->offset = 0;
    if (periodic) {
        double v[OR_MAXDEG + 1];
        int const jmin = bs_eval_basis(b, s->pts[D], v);
        if (D % 2 == 0) s->offset = jmin + D / 2 - D;
        else {
            int const mid = D / 2;
            s->offset = jmin + (v[mid] > v[mid + 1] ? mid : mid + 1) - D;
        }
    }
    /* block sizes, :668-709 (upper from SBCLower, lower from SBCUpper) */
    int upper_bs, lower_bs;
    if (b->uniform) {
        upper_bs = bcl == OR_PERIODIC ? D / 2 : (bcl == OR_GREVILLE ? D - 1 : s->nbe_l);
        lower_bs = bcu == OR_PERIODIC ? D / 2 : (bcu == OR_GREVILLE ? D - 1 : s->nbe_r);
    } else {
        upper_bs = (bcl == OR_PERIODIC || bcl == OR_GREVILLE) ? D - 1 : s->nbe_l + 1;
        lower_bs = (bcu == OR_PERIODIC || bcu == OR_GREVILLE) ? D - 1 : s->nbe_r + 1;
    }
    /* allocate_matrix, :719-774 */
    int const ubw = b->uniform ? D / 2 : D - 1;
    if (periodic) s->mat = maker_periodic_band(nb, ubw, ubw, b->uniform);
    else s->mat = maker_block_band(nb, ubw, ubw, b->uniform, lower_bs, upper_bs);
    /* build_matrix_system, :784-886 */
    double der[(OR_MAXDEG + 1) * (OR_MAXDEG / 2 + 1)];
    if (bcl == OR_HERMITE || bcl == OR_HOMOGENEOUS_HERMITE) {
        int const n = s->nbe_l, ld = n + 1;
        bs_eval_basis_and_n_derivs(b, bs_rmin(b), n, der);
        for (int i = 0; i < D + 1; ++i)
            for (int j = 1; j < n + 1; ++j) {
                double r = 1.0;
                for (int k = 0; k < j; ++k) r *= s->dx;
                der[i * ld + j] *= r;
            }
        for (int i = 0; i < s->nbe_l; ++i)
            for (int j = 0; j < D; ++j) prob_set(s->mat, i, j, der[j * ld + i + s->odd]);
    }
    {
        double v[OR_MAXDEG + 1];
        for (int ix = 0; ix < s->npts; ++ix) {
            int const jmin = bs_eval_basis(b, s->pts[ix], v);
            for (int k = 0; k < D + 1; ++k) {
                int const j = modulo_i(jmin - s->offset + k, nb);
                prob_set(s->mat, ix + s->nbe_l, j, v[k]);
            }
        }
    }
    if (bcu == OR_HERMITE || bcu == OR_HOMOGENEOUS_HERMITE) {
        int const n = s->nbe_r, ld = n + 1;
        bs_eval_basis_and_n_derivs(b, bs_rmax(b), n, der);
        for (int i = 0; i < D + 1; ++i)
            for (int j = 1; j < n + 1; ++j) {
                double r = 1.0;
                for (int k = 0; k < j; ++k) r *= s->dx;
                der[i * ld + j] *= r;
            }
        int const i0 = nb - s->nbe_r, j0 = nb - D;
        for (int j = 0; j < D; ++j)
            for (int i = 0; i < s->nbe_r; ++i) prob_set(s->mat, i0 + i, j0 + j, der[(j + 1) * ld + i + s->odd]);
    }
    int const info = prob_setup(s->mat);
    if (info) {
        snprintf(g_err, sizeof g_err, "LAPACK factorisation failed with error code %d", info);
        or_spline_destroy(s);
        return NULL;
    }
    return s;
}

int or_spline_npts(void* h) { return ((or_spline*)h)->npts; }
void or_spline_points(void* h, double* out) { or_spline* s = (or_spline*)h; memcpy(out, s->pts, sizeof(double) * s->npts); }
int or_spline_nbasis(void* h) { return bs_nbasis(&((or_spline*)h)->bs); }
int or_spline_size(void* h) { return bs_size(&((or_spline*)h)->bs); }
int or_spline_offset(void* h) { return ((or_spline*)h)->offset; }
int or_spline_matrix_type(void* h) { return ((or_spline*)h)->mat->type; }
void* or_spline_problem(void* h) { return ((or_spline*)h)->mat; }
double or_spline_dx(void* h) { return ((or_spline*)h)->dx; }
double or_spline_rmin(void* h) { return bs_rmin(&((or_spline*)h)->bs); }
double or_spline_rmax(void* h) { return bs_rmax(&((or_spline*)h)->bs); }
int or_spline_nknots(void* h) { const or_bspl* b = &((or_spline*)h)->bs; return b->ncells + 1 + 2 * b->degree; }
void or_spline_knots(void* h, double* out)
{
    const or_bspl* b = &((or_spline*)h)->bs;
    for (int k = 0; k < b->ncells + 1 + 2 * b->degree; ++k) out[k] = knot(b, k);
}
int or_spline_eval_basis(void* h, double x, double* v) { return bs_eval_basis(&((or_spline*)h)->bs, x, v); }
int or_spline_eval_basis_and_n_derivs(void* h, double x, int n, double* d)
{
    return bs_eval_basis_and_n_derivs(&((or_spline*)h)->bs, x, n, d);
}
void or_spline_integrals(void* h, double* out) { bs_integrals(&((or_spline*)h)->bs, out); }

/* Layout of a batched array: element (row r, batch line l) lives at
 *   r*rs + (l % inner)*is + (l / inner)*os
 * which covers every layout_right dimension order of 1-, 2- and 3-D chunks. */
typedef struct {
    long rs, inner, is, os;
} or_lay;
static inline long lay_off(const or_lay* L, long r, long l) { return r * L->rs + (l % L->inner) * L->is + (l / L->inner) * L->os; }

/* SplineBuilder::operator(), spline_builder.hpp:896-1101.  dl/dr: Hermite
 * derivatives (nbv rows x batch, layout dlay); NULL when not Hermite. */
int or_spline_build(void* h, const double* vals, long vrs, long vinner, long vis, long vos,
                    double* coef, long crs, long cinner, long cis, long cos_,
                    const double* dl, const double* dr, long drs, long dinner, long dis, long dos, long nbatch)
{
    or_spline* s = (or_spline*)h;
    const or_bspl* b = &s->bs;
    or_lay const VL = {vrs, vinner, vis, vos}, CL = {crs, cinner, cis, cos_}, DLY = {drs, dinner, dis, dos};
    int const nb = bs_nbasis(b), D = b->degree, off = s->offset;
    int const nrows = prob_rhs_rows(s->mat);
    if (nbatch == 0) return 0;
    /* Hermite lower rows, :949-979 */
#pragma omp parallel for schedule(static)
    for (long l = 0; l < nbatch; ++l) {
        if (s->bcl == OR_HERMITE) {
            for (int i = 0; i < s->nbe_l; ++i) {
                double r = 1.0;
                for (int k = 0; k < i + s->odd; ++k) r *= s->dx;
                coef[lay_off(&CL, i, l)] = dl[lay_off(&DLY, i, l)] * r;
            }
        } else if (s->bcl == OR_HOMOGENEOUS_HERMITE) {
            for (int i = 0; i < s->nbe_l; ++i) coef[lay_off(&CL, i, l)] = 0.0;
        }
        for (int i = 0; i < off; ++i) coef[lay_off(&CL, s->nbe_l + i, l)] = 0.0;
        for (int i = 0; i < s->npts; ++i) coef[lay_off(&CL, s->nbe_l + off + i, l)] = vals[lay_off(&VL, i, l)];
        if (s->bcu == OR_HERMITE) {
            for (int i = 0; i < s->nbe_r; ++i) {
                double r = 1.0;
                for (int k = 0; k < i + s->odd; ++k) r *= s->dx;
                coef[lay_off(&CL, nb - s->nbe_r + i, l)] = dr[lay_off(&DLY, i, l)] * r;
            }
        } else if (s->bcu == OR_HOMOGENEOUS_HERMITE) {
            for (int i = 0; i < s->nbe_r; ++i) coef[lay_off(&CL, nb - s->nbe_r + i, l)] = 0.0;
        }
    }
    /* transpose into spline_tr (nrows x nbatch, LayoutRight), :1037-1052 */
    double* tr = (double*)malloc(sizeof(double) * (size_t)nrows * nbatch);
#pragma omp parallel for schedule(static)
    for (long l = 0; l < nbatch; ++l)
        for (int i = 0; i < nb; ++i) tr[(size_t)i * nbatch + l] = coef[lay_off(&CL, i + off, l)];
    /* solve, :1059 */
#pragma omp parallel for schedule(static)
    for (long l = 0; l < nbatch; ++l) prob_solve_col(s->mat, tr + l, nbatch, 0);
    /* transpose back, :1061-1071 */
#pragma omp parallel for schedule(static)
    for (long l = 0; l < nbatch; ++l)
        for (int i = 0; i < nb; ++i) coef[lay_off(&CL, i + off, l)] = tr[(size_t)i * nbatch + l];
    free(tr);
    /* periodic duplication, :1074-1100 */
    if (b->periodic) {
#pragma omp parallel for schedule(static)
        for (long l = 0; l < nbatch; ++l) {
            if (off != 0) {
                for (int i = 0; i < off; ++i) coef[lay_off(&CL, i, l)] = coef[lay_off(&CL, nb + i, l)];
                for (int i = off; i < D; ++i) coef[lay_off(&CL, nb + i, l)] = coef[lay_off(&CL, i, l)];
            }
            for (int i = 0; i < D; ++i) coef[lay_off(&CL, nb + i, l)] = coef[lay_off(&CL, i, l)];
        }
    }
    return 0;
}

/* Adjoint of or_spline_build: given gcoef (N_c x batch), returns gvals
 * (npts x batch) and gdl/gdr (nbv x batch) such that
 * <gcoef, build(vals, dl, dr)> = <gvals, vals> + <gdl, dl> + <gdr, dr>.
 * Built only from the reference's transposed solve primitive
 * (SplinesLinearProblem::solve(b, true), splines_linear_problem.hpp:77). */
int or_spline_build_vjp(void* h, const double* gcoef, long crs, long cinner, long cis, long cos_,
                        double* gvals, long vrs, long vinner, long vis, long vos,
                        double* gdl, double* gdr, long drs, long dinner, long dis, long dos, long nbatch)
{
    or_spline* s = (or_spline*)h;
    const or_bspl* b = &s->bs;
    or_lay const VL = {vrs, vinner, vis, vos}, CL = {crs, cinner, cis, cos_}, DLY = {drs, dinner, dis, dos};
    int const nb = bs_nbasis(b), D = b->degree, off = s->offset, N = bs_size(b);
    int const nrows = prob_rhs_rows(s->mat);
    double* tr = (double*)calloc((size_t)nrows * (nbatch > 0 ? nbatch : 1), sizeof(double));
#pragma omp parallel for schedule(static)
    for (long l = 0; l < nbatch; ++l) {
        /* coef[j] = x[(j - off) mod nb] for j < N (periodic); coef = x otherwise */
        for (int j = 0; j < (b->periodic ? N : nb); ++j) {
            int const k = b->periodic ? modulo_i(j - off, nb) : j;
            tr[(size_t)k * nbatch + l] += gcoef[lay_off(&CL, j, l)];
        }
        (void)D;
    }
#pragma omp parallel for schedule(static)
    for (long l = 0; l < nbatch; ++l) prob_solve_col(s->mat, tr + l, nbatch, 1);
#pragma omp parallel for schedule(static)
    for (long l = 0; l < nbatch; ++l) {
        /* rhs row i+? : rhs[i] = spline[i + off]; spline[nbe_l + off + p] = vals[p] -> rhs row nbe_l + p */
        for (int p = 0; p < s->npts; ++p) gvals[lay_off(&VL, p, l)] = tr[(size_t)(s->nbe_l + p) * nbatch + l];
        if (s->bcl == OR_HERMITE)
            for (int i = 0; i < s->nbe_l; ++i) {
                double r = 1.0;
                for (int k = 0; k < i + s->odd; ++k) r *= s->dx;
                gdl[lay_off(&DLY, i, l)] = tr[(size_t)i * nbatch + l] * r;
            }
        if (s->bcu == OR_HERMITE)
            for (int i = 0; i < s->nbe_r; ++i) {
                double r = 1.0;
                for (int k = 0; k < i + s->odd; ++k) r *= s->dx;
                gdr[lay_off(&DLY, i, l)] = tr[(size_t)(nb - s->nbe_r + i) * nbatch + l] * r;
            }
    }
    free(tr);
    return 0;
}

/* SplineEvaluator::eval / eval_no_bc, spline_evaluator.hpp:529-605, for one
 * point of one line.  extrap rule: 0 null, 1 constant; the constant rule
 * evaluates at its eval_pos (ConstantExtrapolationRule<DimI>::operator(),
 * constant_extrapolation_rule.hpp:52-72): rp[0] for the lower rule, rp[1] for
 * the upper (rp NULL: rmin, rmax).
 * deriv: 0 = value, k > 0 = k-th derivative (no boundary handling, as deriv()). */
static double eval1(const or_spline* s, double x, const double* coef, long stride, int deriv, int el, int eu,
                    const double* rp)
{
    const or_bspl* b = &s->bs;
    int const D = b->degree;
    double v[OR_MAXDEG + 1];
    if (deriv == 0) {
        if (b->periodic) {
            if (x < bs_rmin(b) || x > bs_rmax(b)) x -= floor((x - bs_rmin(b)) / bs_length(b)) * bs_length(b);
        } else {
            if (x < bs_rmin(b)) {
                if (el == 0) return 0.0;
                x = rp ? rp[0] : bs_rmin(b);
            } else if (x > bs_rmax(b)) {
                if (eu == 0) return 0.0;
                x = rp ? rp[1] : bs_rmax(b);
            }
        }
        int const jmin = bs_eval_basis(b, x, v);
        double y = 0.0;
        for (int i = 0; i < D + 1; ++i) y += coef[(jmin + i) * stride] * v[i];
        return y;
    }
    double der[(OR_MAXDEG + 1) * (OR_MAXDEG + 1)];
    int const jmin = bs_eval_basis_and_n_derivs(b, x, deriv, der);
    for (int i = 0; i < D + 1; ++i) v[i] = der[i * (deriv + 1) + deriv];
    double y = 0.0;
    for (int i = 0; i < D + 1; ++i) y += coef[(jmin + i) * stride] * v[i];
    return y;
}

/* batched evaluate/deriv, spline_evaluator.hpp:257-294 and :357-430 */
int or_spline_eval(void* h, const double* coords, double* out, long prs, long pinner, long pis, long pos, long npts,
                   const double* coef, long crs, long cinner, long cis, long cos_, long nbatch, int deriv, int el, int eu,
                   const double* rp)
{
    or_spline* s = (or_spline*)h;
    or_lay const PL = {prs, pinner, pis, pos}, CL = {crs, cinner, cis, cos_};
#pragma omp parallel for schedule(static)
    for (long l = 0; l < nbatch; ++l) {
        const double* c = coef + lay_off(&CL, 0, l);
        for (long i = 0; i < npts; ++i) {
            long const o = lay_off(&PL, i, l);
            out[o] = eval1(s, coords[o], c, crs, deriv, el, eu, rp);
        }
    }
    return 0;
}

/* Adjoint of or_spline_eval (value form): gcoef += w * B(x); gx = w * f'(x)
 * inside the domain (periodic wrap has slope 1), 0 outside (null rule: value
 * 0; constant rule: value independent of x, basis weights taken at eval_pos). */
int or_spline_eval_vjp(void* h, const double* coords, const double* w, double* gx, long prs, long pinner, long pis,
                       long pos, long npts, const double* coef, double* gcoef, long crs, long cinner, long cis,
                       long cos_, long nbatch, int el, int eu, const double* rp)
{
    or_spline* s = (or_spline*)h;
    const or_bspl* b = &s->bs;
    int const D = b->degree;
    or_lay const PL = {prs, pinner, pis, pos}, CL = {crs, cinner, cis, cos_};
#pragma omp parallel for schedule(static)
    for (long l = 0; l < nbatch; ++l) {
        const double* c = coef + lay_off(&CL, 0, l);
        double* gc = gcoef + lay_off(&CL, 0, l);
        for (int j = 0; j < bs_size(b); ++j) gc[j * crs] = 0.0;
        for (long i = 0; i < npts; ++i) {
            long const o = lay_off(&PL, i, l);
            double x = coords[o];
            double const wi = w[o];
            int inside = 1;
            if (b->periodic) {
                if (x < bs_rmin(b) || x > bs_rmax(b)) x -= floor((x - bs_rmin(b)) / bs_length(b)) * bs_length(b);
            } else if (x < bs_rmin(b)) {
                inside = 0;
                if (el == 0) { gx[o] = 0.0; continue; }
                x = rp ? rp[0] : bs_rmin(b);
            } else if (x > bs_rmax(b)) {
                inside = 0;
                if (eu == 0) { gx[o] = 0.0; continue; }
                x = rp ? rp[1] : bs_rmax(b);
            }
            double v[OR_MAXDEG + 1];
            int const jmin = bs_eval_basis(b, x, v);
            for (int k = 0; k < D + 1; ++k) gc[(jmin + k) * crs] += wi * v[k];
            gx[o] = inside ? wi * eval1(s, x, c, crs, 1, el, eu, rp) : 0.0;
        }
    }
    return 0;
}

/* SplineEvaluator::integrate, spline_evaluator.hpp:493-526 */
int or_spline_integrate(void* h, const double* coef, long crs, long cinner, long cis, long cos_, double* out,
                        long ostride, long nbatch)
{
    or_spline* s = (or_spline*)h;
    or_lay const CL = {crs, cinner, cis, cos_};
    int const N = bs_size(&s->bs);
    double* vals = (double*)malloc(sizeof(double) * N);
    bs_integrals(&s->bs, vals);
#pragma omp parallel for schedule(static)
    for (long l = 0; l < nbatch; ++l) {
        double acc = 0;
        for (int i = 0; i < N; ++i) acc += coef[lay_off(&CL, i, l)] * vals[i];
        out[l * ostride] = acc;
    }
    free(vals);
    return 0;
}

/* SplineBuilder::quadrature_coefficients, spline_builder.hpp:1111-1234.
 * out: nbv_l + npts + nbv_r entries (derivs_xmin, values, derivs_xmax). */
int or_spline_quadrature(void* h, double* out)
{
    or_spline* s = (or_spline*)h;
    int const N = bs_size(&s->bs), n = s->mat->n, rows = prob_rhs_rows(s->mat);
    double* ib = (double*)malloc(sizeof(double) * N);
    double* m = (double*)calloc(rows, sizeof(double));
    bs_integrals(&s->bs, ib);
    for (int i = 0; i < n; ++i) m[i] = ib[i];
    prob_solve_col(s->mat, m, 1, 1);
    int const nv = bs_nbasis(&s->bs) - s->nbv_l - s->nbv_r;
    for (int i = 0; i < s->nbv_l; ++i) {
        double r = 1.0;
        for (int k = 0; k < i + s->odd; ++k) r *= s->dx;
        out[i] = m[i] * r;
    }
    for (int i = 0; i < nv; ++i) out[s->nbv_l + i] = m[s->nbv_l + i];
    for (int i = 0; i < s->nbv_r; ++i) {
        double r = 1.0;
        for (int k = 0; k < i + s->odd; ++k) r *= s->dx;
        out[s->nbv_l + nv + i] = m[s->nbv_l + nv + i] * r;
    }
    free(ib);
    free(m);
    return 0;
}

/* clamp of the not-of-interest coordinate, Kokkos::clamp in
 * constant_extrapolation_rule.hpp:168-172 */
static double clampd(double v, double lo, double hi) { return v < lo ? lo : (hi < v ? hi : v); }

/* SplineEvaluator2D::eval / eval_no_bc, spline_evaluator_2d.hpp:679-825, with
 * periodic wrap per dimension, then the rules of dimension 1 and of dimension
 * 2 in that order (rules[0..3]: lower 1, upper 1, lower 2, upper 2; 0 null, 1
 * constant).  A constant rule returns the value at (eval_pos, the other
 * coordinate clamped to [min, max] unless that dimension is periodic)
 * (ConstantExtrapolationRule<DimI, DimNI>::operator(),
 * constant_extrapolation_rule.hpp:157-195).  rp: pos[4], min[4], max[4] per
 * rule (NULL: the guarded bound, and the other dimension's [rmin, rmax]). */
static double eval2(const or_spline* s1, const or_spline* s2, double x1, double x2, const double* coef, long st1,
                    long st2, int d1, int d2, const int* rules, const double* rp)
{
    const or_bspl* b1 = &s1->bs;
    const or_bspl* b2 = &s2->bs;
    if (d1 == 0 && d2 == 0) {
        if (b1->periodic && (x1 < bs_rmin(b1) || x1 > bs_rmax(b1)))
            x1 -= floor((x1 - bs_rmin(b1)) / bs_length(b1)) * bs_length(b1);
        if (b2->periodic && (x2 < bs_rmin(b2) || x2 > bs_rmax(b2)))
            x2 -= floor((x2 - bs_rmin(b2)) / bs_length(b2)) * bs_length(b2);
        int k = -1;
        if (!b1->periodic && x1 < bs_rmin(b1)) k = 0;
        else if (!b1->periodic && x1 > bs_rmax(b1)) k = 1;
        else if (!b2->periodic && x2 < bs_rmin(b2)) k = 2;
        else if (!b2->periodic && x2 > bs_rmax(b2)) k = 3;
        if (k >= 0) {
            if (rules[k] == 0) return 0.0;
            const or_bspl* bi = k < 2 ? b1 : b2; /* dimension of interest */
            const or_bspl* bn = k < 2 ? b2 : b1; /* the other one */
            double const pos = rp ? rp[k] : ((k & 1) ? bs_rmax(bi) : bs_rmin(bi));
            double const lo = rp ? rp[4 + k] : bs_rmin(bn);
            double const hi = rp ? rp[8 + k] : bs_rmax(bn);
            if (k < 2) {
                x1 = pos;
                if (!b2->periodic) x2 = clampd(x2, lo, hi);
            } else {
                x2 = pos;
                if (!b1->periodic) x1 = clampd(x1, lo, hi);
            }
        }
    }
    double v1[OR_MAXDEG + 1], v2[OR_MAXDEG + 1];
    double der[(OR_MAXDEG + 1) * (OR_MAXDEG + 1)];
    int j1, j2;
    if (d1 == 0) j1 = bs_eval_basis(b1, x1, v1);
    else {
        j1 = bs_eval_basis_and_n_derivs(b1, x1, d1, der);
        for (int i = 0; i < b1->degree + 1; ++i) v1[i] = der[i * (d1 + 1) + d1];
    }
    if (d2 == 0) j2 = bs_eval_basis(b2, x2, v2);
    else {
        j2 = bs_eval_basis_and_n_derivs(b2, x2, d2, der);
        for (int i = 0; i < b2->degree + 1; ++i) v2[i] = der[i * (d2 + 1) + d2];
    }
    double y = 0.0;
    for (int i = 0; i < b1->degree + 1; ++i)
        for (int j = 0; j < b2->degree + 1; ++j) y += coef[(j1 + i) * st1 + (j2 + j) * st2] * v1[i] * v2[j];
    return y;
}

/* batched 2D evaluation: coords are interleaved (x1, x2) pairs; point p of
 * batch l at coords[2*(l*npts + p)], out[l*npts + p]; coef of batch l at
 * coef + l*cbs with strides st1, st2. */
int or_spline2d_eval(void* h1, void* h2, const double* coords, double* out, long npts, const double* coef, long cbs,
                     long st1, long st2, long nbatch, int d1, int d2, const int* rules, const double* rp)
{
    or_spline* s1 = (or_spline*)h1;
    or_spline* s2 = (or_spline*)h2;
    long const total = npts * nbatch;
#pragma omp parallel for schedule(static)
    for (long t = 0; t < total; ++t) {
        long const l = t / npts;
        out[t] = eval2(s1, s2, coords[2 * t], coords[2 * t + 1], coef + l * cbs, st1, st2, d1, d2, rules, rp);
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* N-dimensional tensor-product evaluation (SplineEvaluatorND,                */
/* spline_evaluator_nd.hpp; SplineEvaluator3D is its N = 3 case)              */
/* ------------------------------------------------------------------------- */

#define OR_MAXND 6

/* SplineEvaluatorND::eval (spline_evaluator_nd.hpp:684-720) for all orders
 * zero: every periodic dimension is wrapped first (update_coord_eval,
 * :636-652), then the first non-periodic dimension outside [rmin, rmax]
 * decides (check_needs_extrapolation, :655-680): rule 0 (Null) gives 0.
 * With any nonzero order the point goes straight to eval_no_bc, as deriv()
 * does (:445-456).  eval_no_bc (:782-826): per-dimension basis values
 * (get_jmin, :722-751), then y += coef(jmin + idx) * (v_1 * (v_2 * (...)))
 * over the multi-index in row-major order (for_each, :753-768; the basis
 * product is a right fold). */
static double evalnd(or_spline* const* s, int nd, const double* xin, const double* coef, const long* st,
                     const int* orders, const int* rules)
{
    double x[OR_MAXND];
    int any = 0;
    for (int k = 0; k < nd; ++k) {
        x[k] = xin[k];
        any |= orders[k];
    }
    if (!any) {
        for (int k = 0; k < nd; ++k) {
            const or_bspl* b = &s[k]->bs;
            if (b->periodic && (x[k] < bs_rmin(b) || x[k] > bs_rmax(b)))
                x[k] -= floor((x[k] - bs_rmin(b)) / bs_length(b)) * bs_length(b);
        }
        for (int k = 0; k < nd; ++k) {
            const or_bspl* b = &s[k]->bs;
            if (b->periodic) continue;
            if (x[k] < bs_rmin(b) && rules[2 * k] == 0) return 0.0;
            if (x[k] > bs_rmax(b) && rules[2 * k + 1] == 0) return 0.0;
        }
    }
    double v[OR_MAXND][OR_MAXDEG + 1];
    int j[OR_MAXND], n[OR_MAXND];
    double der[(OR_MAXDEG + 1) * (OR_MAXDEG + 1)];
    for (int k = 0; k < nd; ++k) {
        const or_bspl* b = &s[k]->bs;
        n[k] = b->degree + 1;
        if (orders[k] == 0) j[k] = bs_eval_basis(b, x[k], v[k]);
        else {
            j[k] = bs_eval_basis_and_n_derivs(b, x[k], orders[k], der);
            for (int i = 0; i < n[k]; ++i) v[k][i] = der[i * (orders[k] + 1) + orders[k]];
        }
    }
    int idx[OR_MAXND] = {0};
    double y = 0.0;
    for (;;) {
        double p = v[nd - 1][idx[nd - 1]];
        for (int k = nd - 2; k >= 0; --k) p = v[k][idx[k]] * p;
        long off = 0;
        for (int k = 0; k < nd; ++k) off += (long)(j[k] + idx[k]) * st[k];
        y += coef[off] * p;
        int k = nd - 1;
        while (k >= 0 && ++idx[k] == n[k]) idx[k--] = 0;
        if (k < 0) break;
    }
    return y;
}

/* batched ND evaluation: coords interleave the nd coordinates of each point;
 * point p of batch l at coords[nd*(l*npts + p)], out[l*npts + p]; coef of
 * batch l at coef + l*cbs, element (i_1..i_nd) at sum i_k st[k].
 * rules: lower and upper rule of each dimension (0 Null). */
int or_splinend_eval(void* const* hs, int nd, const double* coords, double* out, long npts, const double* coef,
                     long cbs, const long* st, long nbatch, const int* orders, const int* rules)
{
    if (nd < 1 || nd > OR_MAXND) { set_err("dimension out of range"); return -1; }
    or_spline* const* s = (or_spline* const*)hs;
    long const total = npts * nbatch;
#pragma omp parallel for schedule(static)
    for (long t = 0; t < total; ++t) {
        long const l = t / npts;
        out[t] = evalnd(s, nd, coords + nd * t, coef + l * cbs, st, orders, rules);
    }
    return 0;
}

/* SplineEvaluatorND::integrate (spline_evaluator_nd.hpp:597-632): for each
 * batch, sum over the coefficient multi-index (row-major) of
 * coef(i) * (I_1(i_1) * (I_2(i_2) * (...))), I_k the integrals of dimension
 * k's B-splines (integrals.hpp:30-177). */
int or_splinend_integrate(void* const* hs, int nd, double* out, const double* coef, long cbs, const long* st,
                          long nbatch)
{
    if (nd < 1 || nd > OR_MAXND) { set_err("dimension out of range"); return -1; }
    or_spline* const* s = (or_spline* const*)hs;
    double* q[OR_MAXND];
    int n[OR_MAXND];
    for (int k = 0; k < nd; ++k) {
        n[k] = bs_size(&s[k]->bs);
        q[k] = (double*)malloc(sizeof(double) * n[k]);
        bs_integrals(&s[k]->bs, q[k]);
    }
#pragma omp parallel for schedule(static)
    for (long l = 0; l < nbatch; ++l) {
        int idx[OR_MAXND] = {0};
        double acc = 0.0;
        for (;;) {
            double p = q[nd - 1][idx[nd - 1]];
            for (int k = nd - 2; k >= 0; --k) p = q[k][idx[k]] * p;
            long off = l * cbs;
            for (int k = 0; k < nd; ++k) off += (long)idx[k] * st[k];
            acc += coef[off] * p;
            int k = nd - 1;
            while (k >= 0 && ++idx[k] == n[k]) idx[k--] = 0;
            if (k < 0) break;
        }
        out[l] = acc;
    }
    for (int k = 0; k < nd; ++k) free(q[k]);
    return 0;
}
