// Host-side FP64 setup of spline plans: knots, interpolation points, matrix
// assembly, block partition and factorisation.  Runs once per builder; the
// factored matrix is then uploaded as a compact device plan that the batched
// kernels sweep.  Algorithms follow the reference step by step (citations per
// function); the factorisations are the netlib LAPACK algorithms the reference
// calls through LAPACKE (dgbtf2 via dgbtrf, dpbtf2 via dpbtrf, dpttrf, dgetf2).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "internal.hpp"

namespace ddcx {

void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw Error(DDCX_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

DeviceBuffer::~DeviceBuffer()
{
    if (ptr) (void)hipFree(ptr);
}

void DeviceBuffer::upload(const void* host, size_t nbytes)
{
    if (ptr) {
        (void)hipFree(ptr);
        ptr = nullptr;
    }
    bytes = nbytes;
    if (nbytes == 0) return;
    hip_check(hipMalloc(&ptr, nbytes), "hipMalloc");
    hip_check(hipMemcpy(ptr, host, nbytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
}

// ---------------------------------------------------------------------------
// HostQ
// ---------------------------------------------------------------------------

void HostQ::init(QType t, int n_, int kl_, int ku_)
{
    type = t;
    n = n_;
    kl = kl_;
    ku = ku_;
    switch (t) {
    case QType::Dense: q.assign(size_t(n) * n, 0.0); ipiv.assign(n, 0); break;
    case QType::Band: q.assign(size_t(2 * kl + ku + 1) * n, 0.0); ipiv.assign(n, 0); break;
    case QType::PDSBand: q.assign(size_t(kl + 1) * n, 0.0); break;
    case QType::PDSTridiag: q.assign(size_t(2) * n, 0.0); break;
    }
}

// set_element of splines_linear_problem_{dense,band,pds_band,pds_tridiag}.cpp
void HostQ::set(int i, int j, double v)
{
    switch (type) {
    case QType::Dense: q[size_t(i) * n + j] = v; break;
    case QType::Band:
        if (i >= std::max(0, j - ku) && i < std::min(n, j + kl + 1)) q[size_t(kl + ku + i - j) * n + j] = v;
        break;
    case QType::PDSBand:
        if (i > j) std::swap(i, j);
        if (j - i < kl + 1) q[size_t(j - i) * n + i] = v;
        break;
    case QType::PDSTridiag:
        if (i > j) std::swap(i, j);
        if (j - i < 2) q[size_t(j - i) * n + i] = v;
        break;
    }
}

double HostQ::get(int i, int j) const
{
    switch (type) {
    case QType::Dense: return q[size_t(i) * n + j];
    case QType::Band:
        if (i >= std::max(0, j - ku) && i < std::min(n, j + kl + 1)) return q[size_t(kl + ku + i - j) * n + j];
        return 0.0;
    case QType::PDSBand:
        if (i > j) std::swap(i, j);
        return (j - i < kl + 1) ? q[size_t(j - i) * n + i] : 0.0;
    case QType::PDSTridiag:
        if (i > j) std::swap(i, j);
        return (j - i < 2) ? q[size_t(j - i) * n + i] : 0.0;
    }
    return 0.0;
}

namespace {

int factor_pttrf(HostQ& m) // dpttrf
{
    int const n = m.n;
    double* d = m.q.data();
    double* e = m.q.data() + n;
    for (int i = 0; i < n - 1; ++i) {
        if (d[i] <= 0.0) return i + 1;
        double const ei = e[i];
        e[i] = ei / d[i];
        d[i + 1] = d[i + 1] - e[i] * ei;
    }
    if (n > 0 && d[n - 1] <= 0.0) return n;
    return 0;
}

int factor_pbtf2(HostQ& m) // dpbtrf('L') -> dpbtf2 for kd <= 64
{
    int const n = m.n, kd = m.kl;
    auto AB = [&](int r, int c) -> double& { return m.q[size_t(r) * n + c]; };
    for (int j = 0; j < n; ++j) {
        double ajj = AB(0, j);
        if (ajj <= 0.0) return j + 1;
        ajj = std::sqrt(ajj);
        AB(0, j) = ajj;
        int const kn = std::min(kd, n - 1 - j);
        if (kn > 0) {
            double const s = 1.0 / ajj;
            for (int i = 1; i <= kn; ++i) AB(i, j) = s * AB(i, j);
            for (int jj = 0; jj < kn; ++jj) { // dsyr('L', kn, -1, x, trailing)
                double const xj = AB(1 + jj, j);
                if (xj != 0.0) {
                    double const temp = -1.0 * xj;
                    for (int ii = jj; ii < kn; ++ii) AB(ii - jj, j + 1 + jj) = AB(ii - jj, j + 1 + jj) + AB(1 + ii, j) * temp;
                }
            }
        }
    }
    return 0;
}

int factor_gbtf2(HostQ& m) // dgbtrf -> dgbtf2 for ku <= 64
{
    int const n = m.n, kl = m.kl, ku = m.ku, kv = kl + ku;
    auto AB = [&](int r, int c) -> double& { return m.q[size_t(r) * n + c]; };
    int info = 0;
    for (int j = ku + 1; j < std::min(kv, n); ++j)
        for (int i = kv - j; i < kl; ++i) AB(i, j) = 0.0;
    int ju = 0;
    for (int j = 0; j < n; ++j) {
        if (j + kv < n)
            for (int i = 0; i < kl; ++i) AB(i, j + kv) = 0.0;
        int const km = std::min(kl, n - 1 - j);
        int jp = 0;
        double amax = std::fabs(AB(kv, j));
        for (int i = 1; i <= km; ++i)
            if (std::fabs(AB(kv + i, j)) > amax) {
                amax = std::fabs(AB(kv + i, j));
                jp = i;
            }
        m.ipiv[j] = jp + j;
        if (AB(kv + jp, j) != 0.0) {
            ju = std::max(ju, std::min(j + ku + jp, n - 1));
            if (jp != 0)
                for (int c = j; c <= ju; ++c) std::swap(AB(kv + jp + j - c, c), AB(kv + j - c, c));
            if (km > 0) {
                double const s = 1.0 / AB(kv, j);
                for (int i = 1; i <= km; ++i) AB(kv + i, j) = s * AB(kv + i, j);
                for (int c = j + 1; c <= ju; ++c) { // dger(km, ju-j, -1, l, u_row, A)
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
    return info;
}

int factor_getf2(HostQ& m) // dgetrf, unblocked right-looking order
{
    int const n = m.n;
    auto A = [&](int i, int j) -> double& { return m.q[size_t(i) * n + j]; };
    int info = 0;
    for (int j = 0; j < n; ++j) {
        int jp = j;
        double amax = std::fabs(A(j, j));
        for (int i = j + 1; i < n; ++i)
            if (std::fabs(A(i, j)) > amax) {
                amax = std::fabs(A(i, j));
                jp = i;
            }
        m.ipiv[j] = jp;
        if (A(jp, j) != 0.0) {
            if (jp != j)
                for (int c = 0; c < n; ++c) std::swap(A(j, c), A(jp, c));
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
    return info;
}

} // namespace

int HostQ::factor()
{
    switch (type) {
    case QType::Dense: return factor_getf2(*this);
    case QType::Band: return factor_gbtf2(*this);
    case QType::PDSBand: return factor_pbtf2(*this);
    case QType::PDSTridiag: return factor_pttrf(*this);
    }
    return -1;
}

// Solves with the factored block (dgetrs / dgbtrs / dpbtrs / dpttrs order).
void HostQ::solve(double* b, long ld, bool transpose) const
{
    auto B = [&](int i) -> double& { return b[long(i) * ld]; };
    switch (type) {
    case QType::PDSTridiag: {
        const double* d = q.data();
        const double* e = q.data() + n;
        if (n == 1) {
            B(0) = B(0) * (1.0 / d[0]);
            return;
        }
        for (int i = 1; i < n; ++i) B(i) = B(i) - B(i - 1) * e[i - 1];
        B(n - 1) = B(n - 1) / d[n - 1];
        for (int i = n - 2; i >= 0; --i) B(i) = B(i) / d[i] - B(i + 1) * e[i];
        return;
    }
    case QType::PDSBand: {
        int const k = kl;
        auto AB = [&](int r, int c) { return q[size_t(r) * n + c]; };
        for (int j = 0; j < n; ++j) {
            B(j) = B(j) / AB(0, j);
            double const temp = B(j);
            for (int i = j + 1; i <= std::min(n - 1, j + k); ++i) B(i) = B(i) - temp * AB(i - j, j);
        }
        for (int j = n - 1; j >= 0; --j) {
            double temp = B(j);
            for (int i = std::min(n - 1, j + k); i >= j + 1; --i) temp = temp - AB(i - j, j) * B(i);
            B(j) = temp / AB(0, j);
        }
        return;
    }
    case QType::Band: {
        int const kd = kl + ku;
        auto AB = [&](int r, int c) { return q[size_t(r) * n + c]; };
        if (!transpose) {
            for (int j = 0; j < n - 1; ++j) {
                int const lm = std::min(kl, n - 1 - j);
                int const l = ipiv[j];
                if (l != j) std::swap(B(l), B(j));
                double const temp = -1.0 * B(j);
                for (int i = 0; i < lm; ++i) B(j + 1 + i) = B(j + 1 + i) + AB(kd + 1 + i, j) * temp;
            }
            for (int j = n - 1; j >= 0; --j) {
                B(j) = B(j) / AB(kd, j);
                double const temp = B(j);
                for (int i = j - 1; i >= std::max(0, j - kd); --i) B(i) = B(i) - temp * AB(kd + i - j, j);
            }
        } else {
            for (int j = 0; j < n; ++j) {
                double temp = B(j);
                for (int i = std::max(0, j - kd); i <= j - 1; ++i) temp = temp - AB(kd + i - j, j) * B(i);
                B(j) = temp / AB(kd, j);
            }
            for (int j = n - 2; j >= 0; --j) {
                int const lm = std::min(kl, n - 1 - j);
                double temp = 0.0;
                for (int i = 0; i < lm; ++i) temp = temp + AB(kd + 1 + i, j) * B(j + 1 + i);
                B(j) = B(j) + -1.0 * temp;
                int const l = ipiv[j];
                if (l != j) std::swap(B(l), B(j));
            }
        }
        return;
    }
    case QType::Dense: {
        if (n == 0) return;
        auto A = [&](int i, int j) { return q[size_t(i) * n + j]; };
        if (!transpose) {
            for (int j = 0; j < n; ++j)
                if (ipiv[j] != j) std::swap(B(ipiv[j]), B(j));
            for (int j = 0; j < n; ++j)
                for (int i = j + 1; i < n; ++i) B(i) = B(i) - B(j) * A(i, j);
            for (int j = n - 1; j >= 0; --j) {
                B(j) = B(j) / A(j, j);
                for (int i = 0; i < j; ++i) B(i) = B(i) - B(j) * A(i, j);
            }
        } else {
            for (int j = 0; j < n; ++j) {
                double temp = B(j);
                for (int i = 0; i < j; ++i) temp = temp - A(i, j) * B(i);
                B(j) = temp / A(j, j);
            }
            for (int j = n - 1; j >= 0; --j) {
                double temp = B(j);
                for (int i = j + 1; i < n; ++i) temp = temp - A(i, j) * B(i);
                B(j) = temp;
            }
            for (int j = n - 1; j >= 0; --j)
                if (ipiv[j] != j) std::swap(B(ipiv[j]), B(j));
        }
        return;
    }
    }
}

// ---------------------------------------------------------------------------
// HostProblem (2x2 / 3x3 blocks, splines_linear_problem_{2x2,3x3}_blocks.cpp)
// ---------------------------------------------------------------------------

namespace {
// adjust_indices, splines_linear_problem_3x3_blocks.cpp:31-47
inline void adjust3(int top, int nq, int& i, int& j)
{
    if (i < top) i += nq;
    else if (i < top + nq) i -= top;
    if (j < top) j += nq;
    else if (j < top + nq) j -= top;
}
} // namespace

void HostProblem::set(int i, int j, double v)
{
    if (!blocked) return q.set(i, j, v);
    int const nq_ = q.n, nb_ = n - q.n;
    adjust3(top, nq_, i, j);
    if (i < nq_ && j < nq_) q.set(i, j, v);
    else if (i >= nq_ && j >= nq_) schur.set(i - nq_, j - nq_, v);
    else if (j >= nq_) gamma[size_t(i) * nb_ + (j - nq_)] = v;
    else lambda[size_t(i - nq_) * nq_ + j] = v;
}

double HostProblem::get(int i, int j) const
{
    if (!blocked) return q.get(i, j);
    int const nq_ = q.n, nb_ = n - q.n;
    adjust3(top, nq_, i, j);
    if (i < nq_ && j < nq_) return q.get(i, j);
    if (i >= nq_ && j >= nq_) return schur.get(i - nq_, j - nq_);
    if (j >= nq_) return gamma[size_t(i) * nb_ + (j - nq_)];
    return lambda[size_t(i - nq_) * nq_ + j];
}

// SplinesLinearProblem2x2Blocks::setup_solver, 2x2_blocks.cpp:213-234
void HostProblem::setup()
{
    auto check = [](int info, const char* what) {
        if (info != 0) throw Error(DDCX_ERR_SINGULAR, std::string(what) + " failed with error code " + std::to_string(info));
    };
    char const* qname = q.type == QType::Dense ? "LAPACKE_dgetrf"
                        : q.type == QType::Band ? "LAPACKE_dgbtrf"
                        : q.type == QType::PDSBand ? "LAPACKE_dpbtrf"
                                                   : "LAPACKE_dpttrf";
    q_orig = q;
    if (blocked) {
        gamma_orig = gamma;
        delta_orig = schur.q;
    }
    check(q.factor(), qname);
    if (!blocked) return;
    int const nq_ = q.n, nb_ = n - q.n;
    for (int j = 0; j < nb_; ++j) q.solve(gamma.data() + j, nb_, false);
    auto to_coo = [](std::vector<double> const& m, int nr, int nc, std::vector<int>& r, std::vector<int>& c,
                     std::vector<double>& v) {
        r.clear(); c.clear(); v.clear();
        for (int i = 0; i < nr; ++i)
            for (int j = 0; j < nc; ++j) {
                double const a = m[size_t(i) * nc + j];
                if (std::fabs(a) >= 1e-14) {
                    r.push_back(i);
                    c.push_back(j);
                    v.push_back(a);
                }
            }
    };
    to_coo(gamma, nq_, nb_, tr_r, tr_c, tr_v);
    to_coo(lambda, nb_, nq_, bl_r, bl_c, bl_v);
    // compute_schur_complement, 2x2_blocks.cpp:193-210
    for (int i = 0; i < nb_; ++i)
        for (int j = 0; j < nb_; ++j) {
            double val = 0.0;
            for (int l = 0; l < nq_; ++l) val += lambda[size_t(i) * nq_ + l] * gamma[size_t(l) * nb_ + j];
            schur.set(i, j, schur.get(i, j) - val);
        }
    check(schur.factor(), "LAPACKE_dgetrf");
}

// Solve in the 3x3 index space; the 3x3 row interchanges
// (3x3_blocks.cpp:66-128) are applied as an index map, without copies.
void HostProblem::solve(double* b, long ld, bool transpose) const
{
    if (!blocked) return q.solve(b, ld, transpose);
    int const nq_ = q.n, nb_ = n - q.n;
    double* b1 = b + long(top) * ld;
    std::vector<double> b2(nb_);
    auto phys = [&](int k) { return k < top ? k : nq_ + k; }; // border row k -> physical row
    for (int k = 0; k < nb_; ++k) b2[k] = b[long(phys(k)) * ld];
    if (!transpose) {
        q.solve(b1, ld, false);
        for (size_t z = 0; z < bl_v.size(); ++z) b2[bl_r[z]] -= bl_v[z] * b1[long(bl_c[z]) * ld];
        schur.solve(b2.data(), 1, false);
        for (size_t z = 0; z < tr_v.size(); ++z) b1[long(tr_r[z]) * ld] -= tr_v[z] * b2[tr_c[z]];
    } else {
        for (size_t z = 0; z < tr_v.size(); ++z) b2[tr_c[z]] -= tr_v[z] * b1[long(tr_r[z]) * ld];
        schur.solve(b2.data(), 1, true);
        for (size_t z = 0; z < bl_v.size(); ++z) b1[long(bl_c[z]) * ld] -= bl_v[z] * b2[bl_r[z]];
        q.solve(b1, ld, true);
    }
    for (int k = 0; k < nb_; ++k) b[long(phys(k)) * ld] = b2[k];
}

// SplinesLinearProblemMaker, splines_linear_problem_maker.hpp:56-146
void make_band_q(HostQ& q, int n, int kl, int ku, bool pds) // make_new_band
{
    if (kl == ku && kl == 1 && pds) q.init(QType::PDSTridiag, n, 1, 1);
    else if (kl == ku && pds) q.init(QType::PDSBand, n, kl, kl);
    else if (2 * kl + ku + 1 >= n) q.init(QType::Dense, n, 0, 0);
    else q.init(QType::Band, n, kl, ku);
}

static void init_border(HostProblem& pr)
{
    int const nq = pr.q.n, nbd = pr.n - nq;
    pr.schur.init(QType::Dense, nbd, 0, 0);
    pr.gamma.assign(size_t(nq) * nbd, 0.0);
    pr.lambda.assign(size_t(nbd) * nq, 0.0);
}

// make_new_block_matrix_with_band_main_block
void make_block_problem(HostProblem& pr, int n, int kl, int ku, bool pds, int bottom, int top)
{
    if (bottom < 0 || top < 0 || n - top - bottom < 1)
        throw Error(DDCX_ERR_INVALID, "block sizes leave no main block");
    pr.n = n;
    pr.blocked = true;
    pr.top = top;
    make_band_q(pr.q, n - top - bottom, kl, ku, pds);
    init_border(pr);
}

// make_new_periodic_band_matrix
void make_periodic_band_problem(HostProblem& pr, int n, int kl, int ku, bool pds)
{
    if (kl >= n || ku >= n) throw Error(DDCX_ERR_INVALID, "band wider than the matrix");
    int const bottom = std::max(kl, ku), top = n - bottom;
    if (long(bottom) * (n + top) + long(2 * kl + ku + 1) * top >= long(n) * n) {
        pr.n = n;
        pr.blocked = false;
        pr.top = 0;
        pr.q.init(QType::Dense, n, 0, 0);
    } else {
        make_block_problem(pr, n, kl, ku, pds, bottom, 0);
    }
}

// SplinesLinearProblem2x2Blocks(n, Dense(n - bottom)) when top == 0, else
// SplinesLinearProblem3x3Blocks(n, top, Dense(n - top - bottom))
void make_dense_block_problem(HostProblem& pr, int n, int bottom, int top)
{
    if (bottom < 0 || top < 0 || n - top - bottom < 1)
        throw Error(DDCX_ERR_INVALID, "block sizes leave no main block");
    pr.n = n;
    pr.blocked = true;
    pr.top = top;
    pr.q.init(QType::Dense, n - top - bottom, 0, 0);
    init_border(pr);
}

// ---------------------------------------------------------------------------
// B-splines
// ---------------------------------------------------------------------------

std::shared_ptr<BSplinesPlan> make_bsplines(ddcx_bsplines_desc const& d)
{
    if (d.degree < 1 || d.degree > kMaxDegree)
        throw Error(DDCX_ERR_UNSUPPORTED, "degree must be in [1, " + std::to_string(kMaxDegree) + "]");
    if (d.ncells < 1) throw Error(DDCX_ERR_INVALID, "ncells must be positive");
    auto p = std::make_shared<BSplinesPlan>();
    BSplinesView& v = p->view;
    v.degree = d.degree;
    v.uniform = d.uniform ? 1 : 0;
    v.periodic = d.periodic ? 1 : 0;
    v.ncells = d.ncells;
    int const D = d.degree;
    int const nk = d.ncells + 1 + 2 * D;
    p->knots_host.resize(nk);
    if (v.uniform) {
        if (!(d.rmin < d.rmax)) throw Error(DDCX_ERR_INVALID, "rmin must be < rmax");
        // init_ghosted(rmin, rmax, ncells+1, degree, degree), uniform_point_sampling.hpp:163-192
        double const step = (d.rmax - d.rmin) / double(d.ncells);
        v.step = step;
        v.origin = d.rmin - double(D) * step;
        v.inv_step = 1.0 / step;
        for (int k = 0; k < nk; ++k) p->knots_host[k] = v.origin + double(k) * step;
    } else {
        if (!d.breaks) throw Error(DDCX_ERR_INVALID, "non-uniform B-splines need break points");
        std::vector<double>& t = p->knots_host;
        for (int i = 0; i <= d.ncells; ++i) {
            if (i > 0 && !(d.breaks[i] > d.breaks[i - 1]))
                throw Error(DDCX_ERR_INVALID, "Non-uniform break points must be sorted"); // non_uniform_point_sampling.hpp:88-90
            t[D + i] = d.breaks[i];
        }
        double const rmin = t[D], rmax = t[D + d.ncells];
        if (v.periodic) {
            double const period = rmax - rmin;
            for (int i = 1; i < D + 1; ++i) {
                t[D - i] = t[D + d.ncells - i] - period;
                t[D + d.ncells + i] = t[D + i] + period;
            }
        } else {
            for (int i = 1; i < D + 1; ++i) {
                t[D - i] = rmin;
                t[D + d.ncells + i] = rmax;
            }
        }
        v.step = v.origin = v.inv_step = 0.0;
    }
    v.rmin = p->knots_host[D];
    v.rmax = p->knots_host[D + d.ncells];
    if (v.uniform) {
        v.rmin = v.origin + double(D) * v.step;
        v.rmax = v.origin + double(D + d.ncells) * v.step;
    }
    v.length = v.rmax - v.rmin;
    v.knots = p->knots_host.data();
    v.cells = nullptr;
    v.nbucket = 0;
    v.inv_bucket = 0.0;
    if (!v.uniform && d.ncells < 65536) {
        // one bucket per cell on average; bucket k covers
        // [rmin + k w, rmin + (k+1) w), w = (rmax - rmin) / nbucket
        int const nbk = d.ncells;
        double const w = v.length / nbk;
        p->inv_bucket = double(nbk) / v.length;
        p->cells_host.resize(nbk);
        for (int k = 0; k < nbk; ++k) {
            double const e0 = v.rmin + double(k) * w;
            double const e1 = std::min(v.rmin + double(k + 1) * w, v.rmax);
            int const lo = nonuniform_cell(v, p->knots_host.data(), e0) - D;
            int const hi = nonuniform_cell(v, p->knots_host.data(), e1) - D;
            p->cells_host[k] = uint32_t(lo) | (uint32_t(hi) << 16);
        }
    }

    // integrals(), integrals.hpp:30-153
    int const sz = p->size(), nb = p->nbasis();
    p->integrals.assign(sz, 0.0);
    if (v.uniform) {
        if (v.periodic) {
            for (int i = 0; i < nb; ++i) p->integrals[i] = v.step;
        } else {
            for (int i = D; i < sz - D; ++i) p->integrals[i] = v.step;
            dispatch_degree(D, [&]<int DD>() {
                for (int i = 0; i < DD; ++i) {
                    double ev[DD + 2];
                    double off;
                    (void)uniform_cell(v, v.rmin, off);
                    uniform_basis_from_offset<DD + 2>(off, ev);
                    double d_eval = 0.0;
                    for (int k = 0; k < DD + 2; ++k) d_eval += ev[k];
                    double c_eval = 0.0;
                    for (int k = 0; k < DD - i; ++k) c_eval += ev[k];
                    double const edge = v.step * (d_eval - c_eval);
                    p->integrals[i] = edge;
                    p->integrals[sz - 1 - i] = edge;
                }
            });
        }
    } else {
        double const inv_deg = 1.0 / (D + 1);
        for (int i = 0; i < nb; ++i) p->integrals[i] = (p->knots_host[i + D + 1] - p->knots_host[i]) * inv_deg;
    }
    if (device_present()) p->upload();
    return p;
}

bool device_present()
{
    static bool const present = [] {
        int n = 0;
        return hipGetDeviceCount(&n) == hipSuccess && n > 0;
    }();
    return present;
}

void BSplinesPlan::upload()
{
    if (!view.uniform) knots_dev.upload(knots_host);
    if (!cells_host.empty()) cells_dev.upload(cells_host);
    integrals_dev.upload(integrals);
    on_device = true;
}

void BSplinesPlan::ensure_device() const
{
    if (!on_device) throw Error(DDCX_ERR_HIP, "no HIP device was present when the B-splines were created");
}

void SolveBuffers::upload(HostProblem const& pr, SolveView& sv)
{
    sv.qtype = int(pr.q.type);
    sv.nq = pr.q.n;
    sv.kl = pr.q.kl;
    sv.ku = pr.q.ku;
    sv.nb = pr.nb();
    sv.top = pr.top;
    sv.tr_nnz = int(pr.tr_v.size());
    sv.bl_nnz = int(pr.bl_v.size());
    q.upload(pr.q.q);
    ipiv.upload(pr.q.ipiv);
    slu.upload(pr.schur.q);
    sipiv.upload(pr.schur.ipiv);
    trr.upload(pr.tr_r);
    trc.upload(pr.tr_c);
    trv.upload(pr.tr_v);
    blr.upload(pr.bl_r);
    blc.upload(pr.bl_c);
    blv.upload(pr.bl_v);
    sv.q = q.as<double>();
    sv.ipiv = ipiv.as<int>();
    sv.s_lu = slu.as<double>();
    sv.s_ipiv = sipiv.as<int>();
    sv.tr_r = trr.as<int>();
    sv.tr_c = trc.as<int>();
    sv.tr_v = trv.as<double>();
    sv.bl_r = blr.as<int>();
    sv.bl_c = blc.as<int>();
    sv.bl_v = blv.as<double>();
}

void BuilderPlan::ensure_device() const
{
    bs->ensure_device();
    if (!on_device) throw Error(DDCX_ERR_HIP, "no HIP device was present when the builder was created");
}

void BuilderPlan::upload()
{
    {
        d_solve.upload(prob, view.sv);
        d_points.upload(points);
        SolveView& sv = view.sv;
        if (tiled.ok) {
            t_Mf.upload(tiled.Mf);
            t_Mb.upload(tiled.Mb);
            t_rt.upload(tiled.rt);
            t_gseg.upload(tiled.gseg);
            t_Gblk.upload(tiled.Gblk);
            TiledView& tv = tview;
            tv.K = tiled.K;
            tv.L = tiled.L;
            tv.kf = tiled.kf;
            tv.kb = tiled.kb;
            tv.wf = tiled.wf;
            tv.wb = tiled.wb;
            tv.nq = prob.q.n;
            tv.nb = prob.nb();
            tv.top = prob.top;
            tv.q0 = bs->view.periodic ? 0 : prob.top - nbe_l;
            tv.bl_nnz = int(prob.bl_v.size());
            tv.rs = tiled.rs;
            tv.ngb = tv.nb > 0 ? int(tiled.Gblk.size() / (size_t(tiled.L) * tv.nb)) : 0;
            tv.rt = t_rt.as<double>();
            tv.Mf = t_Mf.as<double>();
            tv.Mb = t_Mb.as<double>();
            tv.gseg = t_gseg.as<int>();
            tv.Gblk = t_Gblk.as<double>();
            tv.bl_r = sv.bl_r;
            tv.bl_c = sv.bl_c;
            tv.bl_v = sv.bl_v;
            tv.s_lu = sv.s_lu;
            tv.s_ipiv = sv.s_ipiv;
        }
        auto view_of = [&](TiledHost const& th, DeviceBuffer& mf, DeviceBuffer& mb, DeviceBuffer& rt,
                           DeviceBuffer& gs, DeviceBuffer& gb, TiledView& tv) {
            mf.upload(th.Mf);
            mb.upload(th.Mb);
            rt.upload(th.rt);
            gs.upload(th.gseg);
            gb.upload(th.Gblk);
            tv.K = th.K;
            tv.L = th.L;
            tv.kf = th.kf;
            tv.kb = th.kb;
            tv.wf = th.wf;
            tv.wb = th.wb;
            tv.nq = prob.q.n;
            tv.nb = prob.nb();
            tv.top = prob.top;
            tv.q0 = bs->view.periodic ? 0 : prob.top - nbe_l;
            tv.bl_nnz = int(prob.bl_v.size());
            tv.rs = th.rs;
            tv.ngb = tv.nb > 0 ? int(th.Gblk.size() / (size_t(th.L) * tv.nb)) : 0;
            tv.rt = rt.as<double>();
            tv.Mf = mf.as<double>();
            tv.Mb = mb.as<double>();
            tv.gseg = gs.as<int>();
            tv.Gblk = gb.as<double>();
            tv.bl_r = sv.bl_r;
            tv.bl_c = sv.bl_c;
            tv.bl_v = sv.bl_v;
            tv.s_lu = sv.s_lu;
            tv.s_ipiv = sv.s_ipiv;
        };
        if (tiled_f.ok) view_of(tiled_f, tf_Mf, tf_Mb, tf_rt, tf_gseg, tf_Gblk, tview_f);
        if (tiled_t.ok) {
            tt_Mf.upload(tiled_t.Mf);
            tt_Mb.upload(tiled_t.Mb);
            tt_rt.upload(tiled_t.rt);
            tt_gseg.upload(tiled_t.gseg);
            tt_Gblk.upload(tiled_t.Gblk);
            TiledView& tv = tview_t;
            tv = tview; // sizes and border data are those of the solve
            tv.K = tiled_t.K;
            tv.L = tiled_t.L;
            tv.kf = tiled_t.kf;
            tv.kb = tiled_t.kb;
            tv.wf = tiled_t.wf;
            tv.wb = tiled_t.wb;
            tv.nq = prob.q.n;
            tv.nb = prob.nb();
            tv.top = prob.top;
            tv.q0 = bs->view.periodic ? 0 : prob.top - nbe_l;
            tv.bl_nnz = int(prob.bl_v.size());
            tv.rs = tiled_t.rs;
            tv.ngb = tv.nb > 0 ? int(tiled_t.Gblk.size() / (size_t(tiled_t.L) * tv.nb)) : 0;
            tv.rt = tt_rt.as<double>();
            tv.Mf = tt_Mf.as<double>();
            tv.Mb = tt_Mb.as<double>();
            tv.gseg = tt_gseg.as<int>();
            tv.Gblk = tt_Gblk.as<double>();
            tv.bl_r = sv.bl_r;
            tv.bl_c = sv.bl_c;
            tv.bl_v = sv.bl_v;
            tv.s_lu = sv.s_lu;
            tv.s_ipiv = sv.s_ipiv;
        }
        if (refine.ok) {
            RefineHost const& R = refine;
            r_rt.upload(R.rt32);
            r_Mf.upload(R.Mf32);
            r_Mb.upload(R.Mb32);
            r_gseg_blk.upload(R.t.gseg);
            r_Gblk.upload(R.Gblk32);
            r_rq.upload(R.rq);
            r_gr.upload(R.gam_r);
            r_gc.upload(R.gam_c);
            r_gv.upload(R.gam_v);
            r_gamseg.upload(R.gam_seg);
            r_delta.upload(R.delta);
            RefineView& rv = rview;
            rv.K = R.t.K;
            rv.L = R.t.L;
            rv.kf = R.t.kf;
            rv.kb = R.t.kb;
            rv.wf = R.t.wf;
            rv.wb = R.t.wb;
            rv.nq = prob.q.n;
            rv.nb = prob.nb();
            rv.top = prob.top;
            rv.q0 = bs->view.periodic ? 0 : prob.top - nbe_l;
            rv.bl_nnz = int(prob.bl_v.size());
            rv.rs = R.rs;
            rv.kq = R.kq;
            rv.steps = R.steps;
            rv.shift = R.shift;
            rv.rt = r_rt.as<float>();
            rv.Mf = r_Mf.as<float>();
            rv.Mb = r_Mb.as<float>();
            rv.gseg = r_gseg_blk.as<int>();
            rv.Gblk = r_Gblk.as<float>();
            rv.rq = r_rq.as<double>();
            rv.gam_r = r_gr.as<int>();
            rv.gam_c = r_gc.as<int>();
            rv.gam_v = r_gv.as<double>();
            rv.gam_seg = r_gamseg.as<int>();
            rv.delta = r_delta.as<double>();
            rv.bl_r = sv.bl_r;
            rv.bl_c = sv.bl_c;
            rv.bl_v = sv.bl_v;
            rv.s_lu = sv.s_lu;
            rv.s_ipiv = sv.s_ipiv;
        }
    }
    on_device = true;
}

// ---------------------------------------------------------------------------
// Builder
// ---------------------------------------------------------------------------

namespace {

int n_boundary_equations(ddcx_closure bc, int degree) // spline_builder_closures.hpp:33-55
{
    return (bc == DDCX_HERMITE || bc == DDCX_HOMOGENEOUS_HERMITE) ? degree / 2 : 0;
}

// GrevilleInterpolationPoints<BSplines, BCl, BCu>::get_sampling,
// greville_interpolation_points.hpp:46-269
std::vector<double> greville_points(BSplinesPlan const& bsp, ddcx_closure bcl, ddcx_closure bcu)
{
    BSplinesView const& b = bsp.view;
    int const D = b.degree, nb = bsp.nbasis();
    int const nbe_min = n_boundary_equations(bcl, D), nbe_max = n_boundary_equations(bcu, D);
    int const npoints = nb - nbe_min - nbe_max;
    std::vector<double> out(std::max(npoints, 0));
    bool const uniform_sampling = b.uniform && ((nbe_min != 0 && nbe_max != 0) || b.periodic);
    double const shift = (D % 2 == 0) ? 0.5 : 0.0;
    double const dxg = (b.rmax - b.rmin) / b.ncells;
    auto kn = [&](int k) { return bsp.knots_host[k]; };
    if (uniform_sampling) {
        double const origin = b.rmin + shift * dxg;
        for (int i = 0; i < npoints; ++i) out[i] = origin + double(i) * dxg;
        return out;
    }
    if (b.uniform) {
        double const origin = b.rmin + shift * dxg;
        if (bcl == DDCX_GREVILLE) {
            for (int i = 0; i < D / 2 + 1; ++i) {
                out[i] = (D - i) * b.rmin;
                for (int k = D + 1; k < D + 1 + i; ++k) out[i] += kn(k);
                out[i] /= D;
            }
        } else {
            out[0] = origin + 0.0 * dxg;
        }
        int const n_start = (bcl == DDCX_GREVILLE) ? D / 2 + 1 : 1;
        for (int ip = 1; ip < b.ncells; ++ip) out[ip - 1 + n_start] = origin + double(ip) * dxg;
        if (bcu == DDCX_GREVILLE) {
            for (int i = 0; i < D / 2 + 1; ++i) {
                out[npoints - 1 - i] = (D - i) * b.rmax;
                int const first = (nb - 1 - i) + 1;
                for (int k = first; k < first + i; ++k) out[npoints - 1 - i] += kn(k);
                out[npoints - 1 - i] /= D;
            }
        } else {
            out[npoints - 1] = origin + double(b.ncells - 1 + D % 2) * dxg;
        }
        return out;
    }
    int const ng = b.periodic ? nb + 1 : nb;
    std::vector<double> g(ng + 1);
    for (int ib = 0; ib < nb; ++ib) {
        g[ib] = 0.0;
        for (int k = ib + 1; k < ib + 1 + D; ++k) g[ib] += kn(k);
        g[ib] /= D;
    }
    if (b.periodic) {
        std::vector<double> temp;
        int np = 0;
        while (g[np] < b.rmin) {
            temp.push_back(g[np] + b.length);
            ++np;
        }
        for (int i = 0; i < nb - np; ++i) g[i] = g[i + np];
        for (int i = 0; i < np; ++i) g[nb - np + i] = temp[i];
        g[ng - 1] = g[0] + b.length;
    }
    for (int i = 0; i < npoints; ++i) out[i] = g[nbe_min + i];
    return out;
}

// check_valid_grid, spline_builder.hpp:1273-1317
void check_valid_grid(BSplinesPlan const& bsp, std::vector<double> const& pts, int expected)
{
    if (int(pts.size()) != expected)
        throw Error(DDCX_ERR_GRID, "Incorrect number of points supplied to NonUniformInterpolationPoints. (Received : "
                                       + std::to_string(pts.size()) + ", expected : " + std::to_string(expected));
    int const D = bsp.view.degree;
    auto check = [&](int npc, int cell_end) {
        if (npc > D + 1)
            throw Error(DDCX_ERR_GRID, "The spline problem is overconstrained. There are " + std::to_string(npc)
                                           + " points in the " + std::to_string(cell_end - D) + "-th cell.");
    };
    int npc = 0;
    int cur = D + 1;
    for (double pt : pts) {
        double const edge = bsp.knots_host[cur];
        if (pt > edge) {
            check(npc, cur);
            npc = 1;
            cur += 1;
        } else if (pt == edge) {
            check(npc + 1, cur);
            npc = 1;
            cur += 1;
        } else {
            npc += 1;
        }
    }
    check(npc, cur);
}

int modulo(int x, int y) // math_tools.hpp:56-60
{
    return x - y * int(std::floor(double(x) / y));
}

} // namespace

std::unique_ptr<BuilderPlan> make_builder(std::shared_ptr<BSplinesPlan> bs, ddcx_builder_desc const& d)
{
    auto p = std::make_unique<BuilderPlan>();
    BSplinesView const& b = bs->view;
    int const D = b.degree;
    if ((d.lower == DDCX_PERIODIC) != (d.upper == DDCX_PERIODIC))
        throw Error(DDCX_ERR_INVALID, "Incompatible closure relations");
    if ((d.lower == DDCX_PERIODIC) != bool(b.periodic))
        throw Error(DDCX_ERR_INVALID, "periodic closures require periodic B-splines and vice versa");
    p->bs = bs;
    p->bcl = d.lower;
    p->bcu = d.upper;
    p->solver = d.solver;
    p->precision = d.precision;
    if (d.precision != DDCX_FP64 && d.precision != DDCX_FP32_REFINE) throw Error(DDCX_ERR_INVALID, "unknown precision");
    if (d.refine_steps < 0 || d.refine_steps > 2) throw Error(DDCX_ERR_INVALID, "refine_steps must be 0, 1 or 2");
    p->refine_request = d.refine_steps;
    p->nbe_l = n_boundary_equations(d.lower, D);
    p->nbe_r = n_boundary_equations(d.upper, D);
    p->nbv_l = d.lower == DDCX_HOMOGENEOUS_HERMITE ? 0 : p->nbe_l;
    p->nbv_r = d.upper == DDCX_HOMOGENEOUS_HERMITE ? 0 : p->nbe_r;
    p->odd = D % 2;
    int const nb = bs->nbasis();
    if (d.points) p->points.assign(d.points, d.points + d.npoints);
    else p->points = greville_points(*bs, d.lower, d.upper);
    p->dx = (b.rmax - b.rmin) / b.ncells;
    check_valid_grid(*bs, p->points, nb - p->nbe_l - p->nbe_r);

    // compute_offset, spline_builder.hpp:627-658
    p->offset = 0;
    if (b.periodic) {
        dispatch_degree(D, [&]<int DD>() {
            double v[DD + 1];
            int const jmin = eval_basis<DD>(b, bs->knots_host.data(), p->points[DD], v);
            if (DD % 2 == 0) p->offset = jmin + DD / 2 - DD;
            else {
                int const mid = DD / 2;
                p->offset = jmin + (v[mid] > v[mid + 1] ? mid : mid + 1) - DD;
            }
        });
    }

    // compute_block_sizes_{uniform,non_uniform} (:668-709); the upper block
    // size comes from the lower closure and vice versa (:248-254)
    auto bsize = [&](ddcx_closure bc, int nbc) {
        if (b.uniform) {
            if (bc == DDCX_PERIODIC) return D / 2;
            if (bc == DDCX_GREVILLE) return D - 1;
            return nbc;
        }
        if (bc == DDCX_PERIODIC || bc == DDCX_GREVILLE) return D - 1;
        return nbc + 1;
    };
    int const upper_bs = bsize(d.lower, p->nbe_l);
    int const lower_bs = bsize(d.upper, p->nbe_r);

    // allocate_matrix (:719-774) + SplinesLinearProblemMaker (maker.hpp:56-146)
    int const ubw = b.uniform ? D / 2 : D - 1;
    bool const pds = b.uniform;
    HostProblem& pr = p->prob;
    pr.n = nb;
    if (b.periodic) make_periodic_band_problem(pr, nb, ubw, ubw, pds);
    else make_block_problem(pr, nb, ubw, ubw, pds, lower_bs, upper_bs);

    // build_matrix_system, spline_builder.hpp:784-886
    dispatch_degree(D, [&]<int DD>() {
        const double* kn = bs->knots_host.data();
        auto hermite_rows = [&](double x, int n, auto&& place) {
            constexpr int NMAX = DD / 2;
            double der[(DD + 1) * (NMAX + 1)] = {};
            // eval_basis_and_n_derivs with n derivatives; n <= NMAX
            auto eval_n = [&]<int N>() { eval_basis_and_n_derivs<DD, N>(b, kn, x, der); };
            double tmp[(DD + 1) * (NMAX + 1)] = {};
            switch (n) {
            case 0: eval_n.template operator()<0>(); break;
            case 1: if constexpr (NMAX >= 1) eval_n.template operator()<1>(); break;
            case 2: if constexpr (NMAX >= 2) eval_n.template operator()<2>(); break;
            case 3: if constexpr (NMAX >= 3) eval_n.template operator()<3>(); break;
            default: break;
            }
            // der is (DD+1) x (n+1); scale the j-th derivative by dx^j
            for (int i = 0; i < DD + 1; ++i)
                for (int j = 0; j < n + 1; ++j) tmp[i * (n + 1) + j] = der[i * (n + 1) + j];
            for (int i = 0; i < DD + 1; ++i)
                for (int j = 1; j < n + 1; ++j) {
                    double r = 1.0;
                    for (int k = 0; k < j; ++k) r *= p->dx;
                    tmp[i * (n + 1) + j] *= r;
                }
            place(tmp, n + 1);
        };
        if (d.lower == DDCX_HERMITE || d.lower == DDCX_HOMOGENEOUS_HERMITE) {
            hermite_rows(b.rmin, p->nbe_l, [&](const double* der, int ld) {
                for (int i = 0; i < p->nbe_l; ++i)
                    for (int j = 0; j < DD; ++j) pr.set(i, j, der[j * ld + i + p->odd]);
            });
        }
        double v[DD + 1];
        for (size_t ix = 0; ix < p->points.size(); ++ix) {
            int const jmin = eval_basis<DD>(b, kn, p->points[ix], v);
            for (int s = 0; s < DD + 1; ++s) {
                int const j = modulo(jmin - p->offset + s, nb);
                pr.set(int(ix) + p->nbe_l, j, v[s]);
            }
        }
        if (d.upper == DDCX_HERMITE || d.upper == DDCX_HOMOGENEOUS_HERMITE) {
            hermite_rows(b.rmax, p->nbe_r, [&](const double* der, int ld) {
                int const i0 = nb - p->nbe_r, j0 = nb - DD;
                for (int j = 0; j < DD; ++j)
                    for (int i = 0; i < p->nbe_r; ++i) pr.set(i0 + i, j0 + j, der[(j + 1) * ld + i + p->odd]);
            });
        }
    });
    pr.setup();

    // device plan view (the solve view is filled by upload)
    BuildView& bv = p->view;
    bv.degree = D;
    bv.periodic = b.periodic;
    bv.nbasis = nb;
    bv.size = bs->size();
    bv.offset = p->offset;
    bv.npts = int(p->points.size());
    bv.bcl = int(d.lower);
    bv.bcu = int(d.upper);
    bv.nbe_l = p->nbe_l;
    bv.nbe_r = p->nbe_r;
    bv.odd = p->odd;
    for (int k = 0; k <= kMaxDegree; ++k) {
        double r = 1.0;
        for (int j = 0; j < k; ++j) r *= p->dx;
        bv.dxpow[k] = r;
    }
    make_tiled(*p);
    if (p->precision == DDCX_FP32_REFINE) make_refine(*p);
    if (device_present()) p->upload();
    return p;
}

} // namespace ddcx
