// Host precomputation for the tiled build kernel: row-oriented sweep tables of
// the factored main block (padded to K*L rows), per-row responses to the
// incoming segment state, and the exact segment-carry transfer products.
// All in FP64, once per builder.
#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "internal.hpp"

namespace ddcx {

namespace {

constexpr int kMaxWindow = 8; // segments a carry may cross (about 7 for degree 5 on 8-row segments)
constexpr int kMaxLambda = 64;
constexpr int kMaxBorderTiled = 8;
constexpr double kCarryTol = 0x1p-60;

// Row forms used by the kernel (s[k-1] = y_{i-k} forward, x_{i+k} backward).
double fwd_row(TiledHost const& t, int i, double b, const double* s)
{
    double temp = b;
    for (int k = t.kf; k >= 1; --k) temp = temp - s[k - 1] * t.Lf[size_t(i) * t.kf + (k - 1)];
    return t.fwd_div ? temp * t.rDf[i] : temp;
}

double bwd_row(TiledHost const& t, int i, double y, const double* s)
{
    if (t.bwd_divfirst) return y * t.rDb[i] - s[0] * t.Ub[i];
    double temp = y;
    for (int k = t.kb; k >= 1; --k) temp = temp - s[k - 1] * t.Ub[size_t(i) * t.kb + (k - 1)];
    return temp * t.rDb[i];
}

using Mat = std::vector<double>; // k x k row-major

Mat matmul(Mat const& a, Mat const& b, int k)
{
    Mat c(size_t(k) * k, 0.0);
    for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j) {
            double acc = 0.0;
            for (int l = 0; l < k; ++l) acc += a[size_t(i) * k + l] * b[size_t(l) * k + j];
            c[size_t(i) * k + j] = acc;
        }
    return c;
}

double maxabs(Mat const& a)
{
    double m = 0.0;
    for (double v : a) m = std::max(m, std::fabs(v));
    return m;
}

Mat identity(int k)
{
    Mat m(size_t(k) * k, 0.0);
    for (int i = 0; i < k; ++i) m[size_t(i) * k + i] = 1.0;
    return m;
}

// Reason the tiled kernels cannot take this matrix, or nullptr.
const char* tiled_unsupported(HostProblem const& pr, int min_rows)
{
    int const nq = pr.q.n;
    if (!pr.blocked) return "plain dense matrix";
    if (nq < min_rows) return "main block too small";
    if (nq > 2048) return "main block larger than 2048 rows";
    if (pr.nb() > kMaxBorderTiled) return "border larger than 8";
    if (int(pr.bl_v.size()) > kMaxLambda) return "too many border couplings";
    return nullptr;
}

// Sweep tables of the factored main block on a K x L tiling; the carry
// windows are cut where the transfer products fall below tol.  transpose:
// tables of the solve with Q^T (the PDS blocks are symmetric; for the
// unpivoted band Q^T = U^T L^T is a forward sweep with U^T, scaled by the
// reciprocal pivots, and a unit backward sweep with L^T).
void tile_tables(HostProblem const& pr, int K, int L, double tol, TiledHost& t, bool transpose = false,
                 int min_rows = 256)
{
    t = TiledHost{};
    if (char const* why = tiled_unsupported(pr, min_rows)) { t.why = why; return; }
    HostQ const& q = pr.q;
    int const nq = q.n;
    t.K = K;
    t.L = L;
    int const npad = t.K * t.L;
    switch (q.type) {
    case QType::PDSTridiag:
        t.kf = 1; t.kb = 1; t.fwd_div = 0; t.bwd_divfirst = 1;
        break;
    case QType::PDSBand:
        t.kf = q.kl; t.kb = q.kl; t.fwd_div = 1; t.bwd_divfirst = 0;
        break;
    case QType::Band: {
        for (int j = 0; j < nq; ++j)
            if (q.ipiv[j] != j) { t.why = "pivoted band factor"; return; }
        // without interchanges the U factor keeps the upper bandwidth ku
        // (the sweep uses ku even where the factor's outer diagonals are zero)
        int const kd = q.kl + q.ku;
        for (int c = 0; c < nq; ++c)
            for (int k = q.ku + 1; k <= kd && c - k >= 0; ++k)
                if (q.q[size_t(kd - k) * nq + c] != 0.0) { t.why = "U factor wider than ku"; return; }
        if (transpose) {
            t.kf = q.ku; t.kb = q.kl; t.fwd_div = 1; t.bwd_divfirst = 0;
        } else {
            t.kf = q.kl; t.kb = q.ku; t.fwd_div = 0; t.bwd_divfirst = 0;
        }
        break;
    }
    default: t.why = "dense main block"; return;
    }
    if (t.kf < 1 || t.kb < 1) { t.why = "diagonal main block"; return; }
    if (t.kf > 4 || t.kb > 4) { t.why = "bandwidth above 4"; return; }
    int const kf = t.kf, kb = t.kb;
    t.Lf.assign(size_t(npad) * kf, 0.0);
    t.rDf.assign(npad, 1.0);
    t.Ub.assign(size_t(npad) * kb, 0.0);
    t.rDb.assign(npad, 1.0);
    int const n = nq;
    auto AB = [&](int r, int c) { return q.q[size_t(r) * n + c]; };
    for (int i = 0; i < nq; ++i) {
        switch (q.type) {
        case QType::PDSTridiag: {
            const double* d = q.q.data();
            const double* e = q.q.data() + n;
            if (i >= 1) t.Lf[i] = e[i - 1];
            if (i + 1 < n) t.Ub[i] = e[i];
            t.rDb[i] = 1.0 / d[i];
            break;
        }
        case QType::PDSBand:
            for (int k = 1; k <= kf; ++k)
                if (i - k >= 0) t.Lf[size_t(i) * kf + k - 1] = AB(k, i - k);
            t.rDf[i] = 1.0 / AB(0, i);
            for (int k = 1; k <= kb; ++k)
                if (i + k < n) t.Ub[size_t(i) * kb + k - 1] = AB(k, i);
            t.rDb[i] = 1.0 / AB(0, i);
            break;
        case QType::Band: {
            int const kd = q.kl + q.ku;
            if (transpose) {
                // U(i-k, i) and L(i+k, i) in the band storage
                for (int k = 1; k <= kf; ++k)
                    if (i - k >= 0) t.Lf[size_t(i) * kf + k - 1] = AB(kd - k, i);
                t.rDf[i] = 1.0 / AB(kd, i);
                for (int k = 1; k <= kb; ++k)
                    if (i + k < n) t.Ub[size_t(i) * kb + k - 1] = AB(kd + k, i);
                t.rDb[i] = 1.0;
            } else {
                for (int k = 1; k <= kf; ++k)
                    if (i - k >= 0) t.Lf[size_t(i) * kf + k - 1] = AB(kd + k, i - k);
                for (int k = 1; k <= kb; ++k)
                    if (i + k < n) t.Ub[size_t(i) * kb + k - 1] = AB(kd - k, i + k);
                t.rDb[i] = 1.0 / AB(kd, i);
            }
            break;
        }
        default: break;
        }
    }
    // ---- segment transfer maps (responses of each row to a unit incoming state)
    std::vector<double> gf(size_t(npad) * kf, 0.0), gb(size_t(npad) * kb, 0.0);
    std::vector<Mat> Tf(t.K), Tb(t.K);
    for (int s = 0; s < t.K; ++s) {
        int const r0 = s * t.L;
        Mat T(size_t(kf) * kf, 0.0);
        for (int j = 0; j < kf; ++j) {
            std::vector<double> st(kf, 0.0);
            st[j] = 1.0;
            for (int i = r0; i < r0 + t.L; ++i) {
                double const y = fwd_row(t, i, 0.0, st.data());
                gf[size_t(i) * kf + j] = y;
                for (int k = kf - 1; k > 0; --k) st[k] = st[k - 1];
                st[0] = y;
            }
            for (int k = 0; k < kf; ++k) T[size_t(k) * kf + j] = st[k];
        }
        Tf[s] = T;
        Mat Tbs(size_t(kb) * kb, 0.0);
        for (int j = 0; j < kb; ++j) {
            std::vector<double> st(kb, 0.0);
            st[j] = 1.0;
            for (int i = r0 + t.L - 1; i >= r0; --i) {
                double const x = bwd_row(t, i, 0.0, st.data());
                gb[size_t(i) * kb + j] = x;
                for (int k = kb - 1; k > 0; --k) st[k] = st[k - 1];
                st[0] = x;
            }
            for (int k = 0; k < kb; ++k) Tbs[size_t(k) * kb + j] = st[k];
        }
        Tb[s] = Tbs;
    }
    // ---- window products: in(s) = sum_w P_{s,w} out_loc(s -/+ w)
    auto window = [&](int k, std::vector<Mat> const& T, bool forward, std::vector<double>& M, int& W) -> bool {
        W = 1;
        for (int s = 0; s < t.K; ++s) {
            Mat P = identity(k);
            int w = 1;
            while (true) {
                int const src = forward ? s - w : s + w;
                if (src < 0 || src >= t.K) break;
                P = matmul(P, T[src], k);
                if (maxabs(P) < tol) break;
                ++w;
                if (w > kMaxWindow) return false;
            }
            W = std::max(W, w);
        }
        M.assign(size_t(t.K) * W * k * k, 0.0);
        for (int s = 0; s < t.K; ++s) {
            Mat P = identity(k);
            for (int w = 1; w <= W; ++w) {
                int const src = forward ? s - w : s + w;
                if (src < 0 || src >= t.K) break;
                if (w > 1) P = matmul(P, T[forward ? s - w + 1 : s + w - 1], k);
                std::copy(P.begin(), P.end(), M.begin() + (size_t(s) * W + (w - 1)) * k * k);
            }
        }
        return true;
    };
    if (!window(kf, Tf, true, t.Mf, t.wf)) { t.why = "forward carries decay too slowly"; return; }
    if (!window(kb, Tb, false, t.Mb, t.wb)) { t.why = "backward carries decay too slowly"; return; }
    // ---- Q^-1 gamma by segment (|a| < 1e-14 zeroed, as the reference's COO copy)
    int const nbd = pr.nb();
    t.gseg.assign(t.K, -1);
    t.Gblk.clear();
    for (int s = 0; s < t.K; ++s) {
        bool any = false;
        for (int i = s * t.L; i < std::min((s + 1) * t.L, nq) && !any; ++i)
            for (int c = 0; c < nbd; ++c)
                if (std::fabs(pr.gamma[size_t(i) * nbd + c]) >= 1e-14) any = true;
        if (!any) continue;
        t.gseg[s] = int(t.Gblk.size() / (size_t(t.L) * nbd));
        for (int i = s * t.L; i < (s + 1) * t.L; ++i)
            for (int c = 0; c < nbd; ++c) {
                double const a = i < nq ? pr.gamma[size_t(i) * nbd + c] : 0.0;
                t.Gblk.push_back(std::fabs(a) >= 1e-14 ? a : 0.0);
            }
    }
    // ---- interleaved row table [Lf(kf) | Ub(kb) | rDb | rDf (PDS band only)]
    t.rs = kf + kb + 1 + t.fwd_div;
    t.rt.assign(size_t(npad) * t.rs, 0.0);
    for (int i = 0; i < npad; ++i) {
        double* r = t.rt.data() + size_t(i) * t.rs;
        for (int k = 0; k < kf; ++k) r[k] = t.Lf[size_t(i) * kf + k];
        for (int k = 0; k < kb; ++k) r[kf + k] = t.Ub[size_t(i) * kb + k];
        r[kf + kb] = t.rDb[i];
        if (t.fwd_div) r[kf + kb + 1] = t.rDf[i];
    }
    t.ok = true;
}

} // namespace

void make_tiled(BuilderPlan& p)
{
    // K segments of L rows (K*L >= nq) on 1024 threads, LB = 1024/K lines
    // (32 lines give 256-byte rows in batch-contiguous layouts)
    int const nq = p.prob.q.n;
    if (nq < 256) {
        // short lines (64 to 255 rows): 16 segments of 8 or 9 rows up to 144
        // rows (512-thread workgroups), else 32 segments of 8 rows
        int const K = nq <= 144 ? 16 : 32, L = nq <= 128 || nq > 144 ? 8 : 9;
        tile_tables(p.prob, K, L, kCarryTol, p.tiled, false, 64);
        tile_tables(p.prob, K, L, kCarryTol, p.tiled_t, true, 64);
    } else {
        int const K = nq <= 1024 ? 32 : 64;
        tile_tables(p.prob, K, 32, kCarryTol, p.tiled);
        tile_tables(p.prob, K, 32, kCarryTol, p.tiled_t, true);
    }
    // the fused build + eval kernel keeps 16 (nq <= 1024) or 8 lines per
    // workgroup so that their coefficient tile fits in LDS: 16-row segments
    tile_tables(p.prob, p.prob.q.n <= 1024 ? 64 : 128, 16, kCarryTol, p.tiled_f);
    if (p.tiled_t.ok) {
        int nblk = 0;
        for (int g : p.tiled_t.gseg) nblk += g >= 0;
        if (nblk > kMaxGBlocksVjp) {
            p.tiled_t.ok = false;
            p.tiled_t.why = "border couplings spread over too many segments";
        }
    }
}

namespace {

// Host model of one approximate solve S of the refinement (sequential
// sweeps in FP32 with the FP32 tables, border steps in FP64), in the 2x2
// index space: b1 (nq) and b2 (nb) in, x1 and x2 out.
void approx_solve(RefineHost const& R, HostProblem const& pr, std::vector<double> const& b1,
                  std::vector<double> const& b2, std::vector<double>& x1, std::vector<double>& x2)
{
    TiledHost const& t = R.t;
    int const nq = pr.q.n, nb = pr.nb(), kf = t.kf, kb = t.kb, rs = R.rs;
    auto at = [&](int i, int f) { return R.rt32[size_t(i) * rs + f]; };
    int const oLf = 0, oUb = kf, orDb = kf + kb, orDf = kf + kb + 1;
    auto G = [&](int i, int c) {
        int const b = t.gseg[i / t.L];
        return b < 0 ? 0.0f : R.Gblk32[(size_t(b) * t.L + i % t.L) * nb + c];
    };
    std::vector<float> y(nq);
    for (int i = 0; i < nq; ++i) {
        float tt = float(b1[i]);
        for (int k = kf; k >= 1; --k)
            if (i - k >= 0) tt = std::fma(-y[i - k], at(i, oLf + k - 1), tt);
        if (t.fwd_div) tt = tt * at(i, orDf);
        y[i] = tt;
    }
    for (int i = nq - 1; i >= 0; --i) {
        float tt;
        if (t.bwd_divfirst) {
            tt = y[i] * at(i, orDb);
            if (i + 1 < nq) tt = std::fma(-y[i + 1], at(i, oUb), tt);
        } else {
            tt = y[i];
            for (int k = kb; k >= 1; --k)
                if (i + k < nq) tt = std::fma(-y[i + k], at(i, oUb + k - 1), tt);
            tt = tt * at(i, orDb);
        }
        y[i] = tt;
    }
    x2.assign(b2.begin(), b2.end());
    for (size_t z = 0; z < pr.bl_v.size(); ++z) x2[pr.bl_r[z]] -= pr.bl_v[z] * double(y[pr.bl_c[z]]);
    if (nb > 0) pr.schur.solve(x2.data(), 1, false);
    x1.resize(nq);
    for (int i = 0; i < nq; ++i) {
        float tt = y[i];
        for (int c = 0; c < nb; ++c) tt = std::fma(-G(i, c), float(x2[c]), tt);
        x1[i] = double(tt);
    }
}

// r = b - A x with the assembled matrix, in the 2x2 index space.
void residual(RefineHost const& R, HostProblem const& pr, std::vector<double> const& b1,
              std::vector<double> const& b2, std::vector<double> const& x1, std::vector<double> const& x2,
              std::vector<double>& r1, std::vector<double>& r2)
{
    int const nq = pr.q.n, nb = pr.nb(), kq = R.kq, w = 2 * kq + 1;
    r1.resize(nq);
    for (int i = 0; i < nq; ++i) {
        double tt = b1[i];
        for (int j = 0; j < w; ++j) {
            int const c = i - kq + j;
            if (c >= 0 && c < nq) tt = std::fma(-R.rq[size_t(i) * w + j], x1[c], tt);
        }
        r1[i] = tt;
    }
    for (size_t z = 0; z < R.gam_v.size(); ++z) r1[R.gam_r[z]] -= R.gam_v[z] * x2[R.gam_c[z]];
    r2.assign(b2.begin(), b2.end());
    for (size_t z = 0; z < pr.bl_v.size(); ++z) r2[pr.bl_r[z]] -= pr.bl_v[z] * x1[pr.bl_c[z]];
    for (int k = 0; k < nb; ++k)
        for (int c = 0; c < nb; ++c) r2[k] -= R.delta[size_t(k) * nb + c] * x2[c];
}

} // namespace

void make_refine(BuilderPlan& p)
{
    RefineHost& R = p.refine;
    R = RefineHost{};
    HostProblem const& pr = p.prob;
    // L = 16 rows per thread leaves registers for the FP64 right-hand side
    // and solution next to the FP32 work rows
    int const K = pr.q.n <= 1024 ? 64 : 128;
    tile_tables(pr, K, 16, 0x1p-30, R.t);
    if (!R.t.ok) { R.why = R.t.why; return; }
    TiledHost const& t = R.t;
    HostQ const& q0 = pr.q_orig;
    if (q0.kl != q0.ku || q0.kl != t.kf) { R.why = "unsymmetric band"; return; }
    int const nq = q0.n, nb = pr.nb(), npad = t.K * t.L;
    R.rs = t.rs;
    R.rt32.assign(t.rt.begin(), t.rt.end());
    R.Mf32.assign(t.Mf.begin(), t.Mf.end());
    R.Mb32.assign(t.Mb.begin(), t.Mb.end());
    R.Gblk32.assign(t.Gblk.begin(), t.Gblk.end());
    R.kq = q0.kl;
    int const w = 2 * R.kq + 1;
    R.rq.assign(size_t(npad) * w, 0.0);
    for (int i = 0; i < nq; ++i)
        for (int j = 0; j < w; ++j) {
            int const c = i - R.kq + j;
            if (c >= 0 && c < nq) R.rq[size_t(i) * w + j] = q0.get(i, c);
        }
    for (int i = 0; i < nq; ++i)
        for (int c = 0; c < nb; ++c) {
            double const a = pr.gamma_orig[size_t(i) * nb + c];
            if (a != 0.0) {
                R.gam_r.push_back(i);
                R.gam_c.push_back(c);
                R.gam_v.push_back(a);
            }
        }
    R.gam_seg.assign(t.K + 1, 0);
    for (int s = 0, z = 0; s <= t.K; ++s) {
        while (z < int(R.gam_r.size()) && R.gam_r[z] < s * t.L) ++z;
        R.gam_seg[s] = z;
    }
    R.delta = pr.delta_orig;
    // contraction of the refinement on random right-hand sides
    uint64_t st = 0x9E3779B97F4A7C15ull;
    auto rnd = [&] {
        st ^= st << 13;
        st ^= st >> 7;
        st ^= st << 17;
        return double(int64_t(st >> 11) - (int64_t(1) << 52)) * 0x1p-52;
    };
    std::vector<double> b1(nq), b2(nb), x1, x2, r1, r2, d1, d2;
    double rho = 0.0;
    for (int trial = 0; trial < 3; ++trial) {
        for (double& v : b1) v = rnd();
        for (double& v : b2) v = rnd();
        approx_solve(R, pr, b1, b2, x1, x2);
        residual(R, pr, b1, b2, x1, x2, r1, r2);
        approx_solve(R, pr, r1, r2, d1, d2);
        double num = 0.0, den = 0.0;
        for (int i = 0; i < nq; ++i) {
            num = std::max(num, std::fabs(d1[i]));
            den = std::max(den, std::fabs(x1[i]));
        }
        for (int k = 0; k < nb; ++k) {
            num = std::max(num, std::fabs(d2[k]));
            den = std::max(den, std::fabs(x2[k]));
        }
        rho = std::max(rho, den > 0.0 ? num / den : 1.0);
    }
    R.rho = rho;
    R.shift = rho > 0.0 ? std::clamp(int(std::floor(-std::log2(rho))), 0, 60) : 60;
    R.steps = 0;
    for (int k = 1; k <= 2; ++k)
        if (std::pow(2.0 * rho, k + 1) <= 1e-12) {
            R.steps = k;
            break;
        }
    if (R.steps == 0) { R.why = "FP32 sweeps do not contract enough for 2 refinement steps"; return; }
    if (p.refine_request > 0) R.steps = std::max(R.steps, p.refine_request);
    R.ok = true;
}

} // namespace ddcx
