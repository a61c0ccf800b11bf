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

constexpr int kMaxWindow = 4;
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

} // namespace

void make_tiled(BuilderPlan& p)
{
    TiledHost& t = p.tiled;
    t = TiledHost{};
    HostProblem const& pr = p.prob;
    HostQ const& q = pr.q;
    int const nq = q.n;
    if (!pr.blocked) { t.why = "plain dense matrix"; return; }
    if (nq < 256) { t.why = "main block too small"; return; }
    if (nq > 2048) { t.why = "main block larger than 2048 rows"; return; }
    if (pr.nb() > kMaxBorderTiled) { t.why = "border larger than 8"; return; }
    if (int(pr.bl_v.size()) > kMaxLambda) { t.why = "too many border couplings"; return; }
    // tile shape: K segments of L rows (K*L >= nq), 1024 threads, LB = 1024/K
    // lines (32 lines give 256-byte rows in batch-contiguous layouts)
    if (nq <= 1024) {
        t.K = 32;
        t.L = 32;
    } else {
        t.K = 64;
        t.L = 32;
    }
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
        int const kd = q.kl + q.ku;
        int kbe = 0;
        for (int c = 0; c < nq; ++c)
            for (int k = 1; k <= kd && c - k >= 0; ++k)
                if (q.q[size_t(kd - k) * nq + c] != 0.0) kbe = std::max(kbe, k);
        t.kf = q.kl; t.kb = kbe; t.fwd_div = 0; t.bwd_divfirst = 0;
        break;
    }
    default: t.why = "dense main block"; return;
    }
    if (t.kf < 1 || t.kb < 1) { t.why = "diagonal main block"; return; }
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
            for (int k = 1; k <= kf; ++k)
                if (i - k >= 0) t.Lf[size_t(i) * kf + k - 1] = AB(kd + k, i - k);
            for (int k = 1; k <= kb; ++k)
                if (i + k < n) t.Ub[size_t(i) * kb + k - 1] = AB(kd - k, i + k);
            t.rDb[i] = 1.0 / AB(kd, i);
            break;
        }
        default: break;
        }
    }
    // ---- per-row responses to the incoming state and segment transfer maps
    t.gf.assign(size_t(npad) * kf, 0.0);
    t.gb.assign(size_t(npad) * kb, 0.0);
    std::vector<Mat> Tf(t.K), Tb(t.K);
    for (int s = 0; s < t.K; ++s) {
        int const r0 = s * t.L;
        Mat T(size_t(kf) * kf, 0.0);
        for (int j = 0; j < kf; ++j) {
            std::vector<double> st(kf, 0.0);
            st[j] = 1.0;
            for (int i = r0; i < r0 + t.L; ++i) {
                double const y = fwd_row(t, i, 0.0, st.data());
                t.gf[size_t(i) * kf + j] = y;
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
                t.gb[size_t(i) * kb + j] = x;
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
                if (maxabs(P) < kCarryTol) break;
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
    // ---- border table
    int const nbd = pr.nb();
    t.G.assign(size_t(npad) * std::max(nbd, 1), 0.0);
    for (int i = 0; i < nq; ++i)
        for (int c = 0; c < nbd; ++c) {
            double const a = pr.gamma[size_t(i) * nbd + c];
            t.G[size_t(i) * nbd + c] = std::fabs(a) >= 1e-14 ? a : 0.0;
        }
    // ---- interleaved row table
    t.rs = 2 * kf + 2 * kb + 2 + nbd;
    t.rs += t.rs & 1; // 16-byte aligned rows
    t.rt.assign(size_t(npad) * t.rs, 0.0);
    for (int i = 0; i < npad; ++i) {
        double* r = t.rt.data() + size_t(i) * t.rs;
        for (int k = 0; k < kf; ++k) r[k] = t.Lf[size_t(i) * kf + k];
        for (int k = 0; k < kf; ++k) r[kf + k] = t.gf[size_t(i) * kf + k];
        for (int k = 0; k < kb; ++k) r[2 * kf + k] = t.Ub[size_t(i) * kb + k];
        r[2 * kf + kb] = t.rDb[i];
        for (int k = 0; k < kb; ++k) r[2 * kf + kb + 1 + k] = t.gb[size_t(i) * kb + k];
        r[2 * kf + 2 * kb + 1] = t.rDf[i];
        for (int c = 0; c < nbd; ++c) r[2 * kf + 2 * kb + 2 + c] = t.G[size_t(i) * nbd + c];
    }
    t.ok = true;
}

} // namespace ddcx
