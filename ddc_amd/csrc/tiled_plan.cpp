// Host precomputation for the tiled build kernel: row-oriented sweep tables of
// the factored main block, padded to K*L rows, and the exact segment-carry
// transfer products.  All in FP64, once per builder.
#include <algorithm>
#include <cmath>

#include "internal.hpp"

namespace ddcx {

namespace {

constexpr int kMaxWindow = 4;
constexpr int kMaxLambda = 64;
constexpr int kMaxBorderTiled = 8;
constexpr double kCarryTol = 0x1p-60;

// One forward step of the row form on a state vector s (s[k-1] = y_{i-k}).
double fwd_row(TiledHost const& t, int i, double b, const double* s)
{
    double temp = b;
    for (int k = t.kf; k >= 1; --k) temp = temp - s[k - 1] * t.Lf[size_t(i) * t.kf + (k - 1)];
    return t.fwd_div ? temp / t.Df[i] : temp;
}

// One backward step of the row form (s[k-1] = x_{i+k}).
double bwd_row(TiledHost const& t, int i, double y, const double* s)
{
    if (t.bwd_divfirst) return y / t.Db[i] - s[0] * t.Ub[size_t(i) * t.kb];
    double temp = y;
    for (int k = t.kb; k >= 1; --k) temp = temp - s[k - 1] * t.Ub[size_t(i) * t.kb + (k - 1)];
    return temp / t.Db[i];
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
    t.K = 64;
    t.L = nq <= 1024 ? 16 : 32;
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
    int const kf = std::max(t.kf, 1), kb = std::max(t.kb, 1);
    t.Lf.assign(size_t(npad) * kf, 0.0);
    t.Df.assign(npad, 1.0);
    t.Ub.assign(size_t(npad) * kb, 0.0);
    t.Db.assign(npad, 1.0);
    int const n = nq;
    auto AB = [&](int r, int c) { return q.q[size_t(r) * n + c]; };
    for (int i = 0; i < nq; ++i) {
        switch (q.type) {
        case QType::PDSTridiag: {
            const double* d = q.q.data();
            const double* e = q.q.data() + n;
            if (i >= 1) t.Lf[i] = e[i - 1];
            if (i + 1 < n) t.Ub[i] = e[i];
            t.Db[i] = d[i];
            break;
        }
        case QType::PDSBand:
            for (int k = 1; k <= t.kf; ++k)
                if (i - k >= 0) t.Lf[size_t(i) * kf + k - 1] = AB(k, i - k);
            t.Df[i] = AB(0, i);
            for (int k = 1; k <= t.kb; ++k)
                if (i + k < n) t.Ub[size_t(i) * kb + k - 1] = AB(k, i);
            t.Db[i] = AB(0, i);
            break;
        case QType::Band: {
            int const kd = q.kl + q.ku;
            for (int k = 1; k <= t.kf; ++k)
                if (i - k >= 0) t.Lf[size_t(i) * kf + k - 1] = AB(kd + k, i - k);
            for (int k = 1; k <= t.kb; ++k)
                if (i + k < n) t.Ub[size_t(i) * kb + k - 1] = AB(kd - k, i + k);
            t.Db[i] = AB(kd, i);
            break;
        }
        default: break;
        }
    }
    // ---- transfer matrices of each segment (zero right-hand side)
    std::vector<Mat> Tf(t.K), Tb(t.K);
    for (int s = 0; s < t.K; ++s) {
        int const r0 = s * t.L;
        if (t.kf > 0) {
            Mat T(size_t(t.kf) * t.kf, 0.0);
            for (int j = 0; j < t.kf; ++j) {
                std::vector<double> st(t.kf, 0.0);
                st[j] = 1.0;
                for (int i = r0; i < r0 + t.L; ++i) {
                    double const y = fwd_row(t, i, 0.0, st.data());
                    for (int k = t.kf - 1; k > 0; --k) st[k] = st[k - 1];
                    st[0] = y;
                }
                for (int k = 0; k < t.kf; ++k) T[size_t(k) * t.kf + j] = st[k];
            }
            Tf[s] = T;
        }
        if (t.kb > 0) {
            Mat T(size_t(t.kb) * t.kb, 0.0);
            for (int j = 0; j < t.kb; ++j) {
                std::vector<double> st(t.kb, 0.0);
                st[j] = 1.0;
                for (int i = r0 + t.L - 1; i >= r0; --i) {
                    double const x = bwd_row(t, i, 0.0, st.data());
                    for (int k = t.kb - 1; k > 0; --k) st[k] = st[k - 1];
                    st[0] = x;
                }
                for (int k = 0; k < t.kb; ++k) T[size_t(k) * t.kb + j] = st[k];
            }
            Tb[s] = T;
        }
    }
    // ---- window products and window width
    auto window = [&](int k, std::vector<Mat> const& T, bool forward, std::vector<double>& M, int& W) -> bool {
        if (k == 0) {
            W = 0;
            M.assign(1, 0.0);
            return true;
        }
        W = 1;
        for (int s = 0; s < t.K; ++s) {
            Mat P = identity(k);
            int w = 1;
            while (true) {
                int const src = forward ? s - w : s + w; // segment whose state would enter next
                if (src < 0 || src >= t.K) break;
                P = forward ? matmul(P, T[s - w], k) : matmul(P, T[s + w], k);
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
                if (w > 1) P = forward ? matmul(P, T[s - w + 1], k) : matmul(P, T[s + w - 1], k);
                int const src = forward ? s - w : s + w;
                if (src < 0 || src >= t.K) break;
                std::copy(P.begin(), P.end(), M.begin() + (size_t(s) * W + (w - 1)) * k * k);
            }
        }
        return true;
    };
    if (!window(t.kf, Tf, true, t.Mf, t.wf)) { t.why = "forward carries decay too slowly"; return; }
    if (!window(t.kb, Tb, false, t.Mb, t.wb)) { t.why = "backward carries decay too slowly"; return; }
    // ---- border tables
    int const nbd = pr.nb();
    t.G.assign(size_t(npad) * std::max(nbd, 1), 0.0);
    for (int i = 0; i < nq; ++i)
        for (int c = 0; c < nbd; ++c) {
            double const a = pr.gamma[size_t(i) * nbd + c];
            t.G[size_t(i) * nbd + c] = std::fabs(a) >= 1e-14 ? a : 0.0;
        }
    t.ok = true;
}

} // namespace ddcx
