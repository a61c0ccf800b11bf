// Drives device arrays through the C++ facade (include/ddcx.hpp) the way a
// caller of the reference's SplineBuilder / SplineEvaluator /
// SplineBuilder2D / SplineEvaluator2D would, and writes inputs and outputs
// as raw little-endian doubles into the directory given as argv[1]; the
// test test_gpu_cpp_facade.py checks them against the oracle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "ddcx.hpp"
#include "ddcx_device.hpp"

namespace {

std::string g_dir;

void hip_ok(hipError_t e, const char* what)
{
    if (e != hipSuccess) {
        std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
        std::exit(3);
    }
}

struct DeviceArray {
    double* p = nullptr;
    size_t n = 0;
    explicit DeviceArray(size_t count) : n(count) { hip_ok(hipMalloc(&p, n * sizeof(double)), "hipMalloc"); }
    explicit DeviceArray(std::vector<double> const& h) : DeviceArray(h.size())
    {
        hip_ok(hipMemcpy(p, h.data(), n * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy H2D");
    }
    ~DeviceArray() { (void)hipFree(p); }
    DeviceArray(DeviceArray const&) = delete;
    std::vector<double> get() const
    {
        std::vector<double> h(n);
        hip_ok(hipDeviceSynchronize(), "sync");
        hip_ok(hipMemcpy(h.data(), p, n * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy D2H");
        return h;
    }
};

void dump(std::string const& name, std::vector<double> const& v)
{
    std::FILE* f = std::fopen((g_dir + "/" + name + ".bin").c_str(), "wb");
    if (!f || std::fwrite(v.data(), sizeof(double), v.size(), f) != v.size()) {
        std::fprintf(stderr, "cannot write %s\n", name.c_str());
        std::exit(4);
    }
    std::fclose(f);
}

// deterministic uniform [0, 1) numbers (64-bit LCG, top 53 bits)
struct Lcg {
    uint64_t s;
    double next()
    {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return double(s >> 11) * 0x1p-53;
    }
};

// The caller's own kernel, the way a reference user calls the device-callable
// SplineEvaluator::operator()(coord, coef_1d) / deriv(order, coord, coef_1d)
// (spline_evaluator.hpp:228-235): one thread per point of a batch-contiguous
// (P, B) array, each line's coefficients at stride B.
__global__ void user_point_kernel(ddcx_device_evaluator de, int order, const double* coef, const double* x,
                                  double* out, int P, int B)
{
    int const t = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= P * B) return;
    int const l = t % B;
    ddcx::device::Evaluator const e = ddcx::device::evaluator(de);
    out[t] = ddcx::device::deriv(e, order, x[t], coef + l, B);
}

void point_eval(ddcx::SplineEvaluator const& ev, int order, DeviceArray const& dc, DeviceArray const& dx,
                DeviceArray& out, int P, int B)
{
    ddcx_device_evaluator const de = ev.device_view();
    hipLaunchKernelGGL(user_point_kernel, dim3(unsigned((P * B + 255) / 256)), dim3(256), 0, 0, de, order, dc.p,
                       dx.p, out.p, P, B);
    hip_ok(hipGetLastError(), "user_point_kernel launch");
    hip_ok(hipDeviceSynchronize(), "user_point_kernel");
}

// 1D: uniform periodic cubic, 64 cells, 33 lines, in one layout; build, eval,
// deriv(1), integrate
void case_1d(bool batch_contiguous)
{
    int const nc = 64, B = 33, P = 100;
    ddcx::UniformBSplines bs(3, 0.0, 1.0, nc, true);
    ddcx::SplineBuilder builder(bs, ddcx::SplineBuilderClosure::PERIODIC, ddcx::SplineBuilderClosure::PERIODIC);
    ddcx::SplineEvaluator ev(bs, ddcx::PeriodicExtrapolationRule(), ddcx::PeriodicExtrapolationRule());
    int const m = builder.npoints(), s = bs.size();
    std::vector<double> x = builder.interpolation_points();
    Lcg r{uint64_t(batch_contiguous ? 1 : 2)};
    std::vector<double> vals(size_t(m) * B), coords(size_t(P) * B);
    for (int i = 0; i < m; ++i)
        for (int l = 0; l < B; ++l) {
            double const v = std::sin(2 * M_PI * x[i] * (1 + l % 8)) + 0.1 * (r.next() - 0.5);
            vals[batch_contiguous ? size_t(i) * B + l : size_t(l) * m + i] = v;
        }
    for (auto& c : coords) c = -0.5 + 2.0 * r.next();
    int const ax = batch_contiguous ? 0 : 1;
    std::vector<int64_t> const ev_ext = batch_contiguous ? std::vector<int64_t>{m, B} : std::vector<int64_t>{B, m};
    std::vector<int64_t> const ec_ext = batch_contiguous ? std::vector<int64_t>{s, B} : std::vector<int64_t>{B, s};
    std::vector<int64_t> const ep_ext = batch_contiguous ? std::vector<int64_t>{P, B} : std::vector<int64_t>{B, P};
    int64_t nb = 0;
    ddcx_layout const lv = ddcx::batched_layout(ev_ext, ax, &nb);
    ddcx_layout const lc = ddcx::batched_layout(ec_ext, ax);
    ddcx_layout const lp = ddcx::batched_layout(ep_ext, ax);
    DeviceArray dv(vals), dc(size_t(s) * B), dx(coords), dout(size_t(P) * B), dder(size_t(P) * B), dint(B);
    builder(ddcx::Span<double>{dc.p, lc, nb}, ddcx::Span<double const>{dv.p, lv, nb});
    ev(ddcx::Span<double>{dout.p, lp, nb}, ddcx::Span<double const>{dx.p, lp, nb},
       ddcx::Span<double const>{dc.p, lc, nb}, P);
    ev.deriv(1, ddcx::Span<double>{dder.p, lp, nb}, ddcx::Span<double const>{dx.p, lp, nb},
             ddcx::Span<double const>{dc.p, lc, nb}, P);
    ev.integrate(dint.p, 1, ddcx::Span<double const>{dc.p, lc, nb});
    std::string const tag = batch_contiguous ? "1d_bc_" : "1d_ic_";
    dump(tag + "vals", vals);
    dump(tag + "coords", coords);
    dump(tag + "coef", dc.get());
    dump(tag + "eval", dout.get());
    dump(tag + "deriv1", dder.get());
    dump(tag + "integrals", dint.get());
    if (batch_contiguous) {
        DeviceArray dp0(size_t(P) * B), dp1(size_t(P) * B);
        point_eval(ev, 0, dc, dx, dp0, P, B);
        point_eval(ev, 1, dc, dx, dp1, P, B);
        dump(tag + "point_eval", dp0.get());
        dump(tag + "point_deriv1", dp1.get());
    }
}

// 1D: non-uniform cubic Greville with ConstantExtrapolationRule(eval_pos)
// on both sides, points outside the domain
void case_rules()
{
    int const nc = 40, B = 5, P = 200;
    Lcg r{3};
    std::vector<double> br(nc + 1);
    br[0] = 0.0;
    for (int i = 1; i <= nc; ++i) br[i] = br[i - 1] + 0.5 + r.next();
    for (auto& b : br) b /= br[nc];
    ddcx::NonUniformBSplines bs(3, br, false);
    ddcx::SplineBuilder builder(bs, ddcx::SplineBuilderClosure::GREVILLE, ddcx::SplineBuilderClosure::GREVILLE);
    ddcx::SplineEvaluator ev(bs, ddcx::ConstantExtrapolationRule(0.3), ddcx::ConstantExtrapolationRule(0.8));
    int const m = builder.npoints(), s = bs.size();
    std::vector<double> vals(size_t(m) * B), coords(size_t(P) * B);
    for (auto& v : vals) v = r.next() - 0.5;
    for (auto& c : coords) c = -0.3 + 1.6 * r.next();
    int64_t nb = 0;
    ddcx_layout const lv = ddcx::batched_layout({m, B}, 0, &nb);
    ddcx_layout const lc = ddcx::batched_layout({s, B}, 0);
    ddcx_layout const lp = ddcx::batched_layout({P, B}, 0);
    DeviceArray dv(vals), dc(size_t(s) * B), dx(coords), dout(size_t(P) * B);
    builder(ddcx::Span<double>{dc.p, lc, nb}, ddcx::Span<double const>{dv.p, lv, nb});
    ev(ddcx::Span<double>{dout.p, lp, nb}, ddcx::Span<double const>{dx.p, lp, nb},
       ddcx::Span<double const>{dc.p, lc, nb}, P);
    dump("rules_breaks", br);
    dump("rules_vals", vals);
    dump("rules_coords", coords);
    dump("rules_eval", dout.get());
    DeviceArray dp0(size_t(P) * B), dp2(size_t(P) * B);
    point_eval(ev, 0, dc, dx, dp0, P, B);
    point_eval(ev, 2, dc, dx, dp2, P, B);
    dump("rules_point_eval", dp0.get());
    dump("rules_point_deriv2", dp2.get());
}

// 2D: non-uniform cubic Greville (20 cells) x uniform periodic cubic (16
// cells), 2 batch entries: SplineBuilder2D, SplineEvaluator2D values,
// deriv_dim_1, the packed evaluation and the adjoint
void case_2d()
{
    int const nc1 = 20, nc2 = 16, NB = 2, P = 500;
    Lcg r{4};
    std::vector<double> br(nc1 + 1);
    br[0] = 0.0;
    for (int i = 1; i <= nc1; ++i) br[i] = br[i - 1] + 0.5 + r.next();
    for (auto& b : br) b /= br[nc1];
    ddcx::NonUniformBSplines bs1(3, br, false);
    ddcx::UniformBSplines bs2(3, 0.0, 1.0, nc2, true);
    ddcx::SplineBuilder b1(bs1, ddcx::SplineBuilderClosure::GREVILLE, ddcx::SplineBuilderClosure::GREVILLE);
    ddcx::SplineBuilder b2(bs2, ddcx::SplineBuilderClosure::PERIODIC, ddcx::SplineBuilderClosure::PERIODIC);
    ddcx::SplineEvaluator e1(bs1, ddcx::NullExtrapolationRule(), ddcx::NullExtrapolationRule());
    ddcx::SplineEvaluator e2(bs2, ddcx::PeriodicExtrapolationRule(), ddcx::PeriodicExtrapolationRule());
    ddcx::SplineBuilder2D builder(b1, b2, bs1.size());
    ddcx::SplineEvaluator2D ev(e1, e2);
    int const n1 = b1.npoints(), n2 = b2.npoints(), s1 = bs1.size(), s2 = bs2.size();
    std::vector<double> vals(size_t(NB) * n1 * n2), coords(size_t(NB) * P * 2), w(size_t(NB) * P);
    for (auto& v : vals) v = r.next() - 0.5;
    for (size_t i = 0; i < coords.size(); i += 2) {
        coords[i] = 0.01 + 0.98 * r.next();
        coords[i + 1] = -0.5 + 2.0 * r.next();
    }
    for (auto& x : w) x = r.next() - 0.5;
    DeviceArray dv(vals), dc(size_t(NB) * s1 * s2), dws(size_t(builder.workspace_size(NB))), dx(coords), dw(w);
    DeviceArray dout(size_t(NB) * P), dd1(size_t(NB) * P), dgx(size_t(NB) * P * 2), dgc(size_t(NB) * s1 * s2);
    builder(dc.p, dv.p, NB, dws.p);
    int64_t const cbs = int64_t(s1) * s2;
    ev(dout.p, dx.p, P, dc.p, cbs, s2, 1, NB);
    ev.deriv_dim_1(dd1.p, dx.p, P, dc.p, cbs, s2, 1, NB);
    // the packed form with a workspace of one batch entry (entries packed one at a time)
    DeviceArray dpk(size_t(NB) * P), dpw(size_t(ev.eval_workspace(1) + 7) / 8);
    ev.deriv(0, 0, dpk.p, dx.p, P, dc.p, cbs, s2, 1, NB, dpw.p, ev.eval_workspace(1), nullptr);
    int64_t const wb = ev.eval_vjp_workspace(P, NB);
    DeviceArray dwork(size_t(std::max<int64_t>(wb, 8) + 7) / 8);
    ev.eval_vjp(dx.p, dw.p, dgx.p, P, dc.p, dgc.p, cbs, s2, 1, NB, wb > 0 ? dwork.p : nullptr, wb);
    dump("2d_breaks", br);
    dump("2d_vals", vals);
    dump("2d_coords", coords);
    dump("2d_w", w);
    dump("2d_coef", dc.get());
    dump("2d_eval", dout.get());
    dump("2d_deriv1", dd1.get());
    dump("2d_eval_packed", dpk.get());
    dump("2d_gx", dgx.get());
    dump("2d_gcoef", dgc.get());
}

// 2D with Hermite closures in both dimensions: non-uniform cubic (12 cells)
// x uniform quintic (10 cells), 2 batch entries; SplineBuilder2D with the
// boundary and corner derivatives, and its adjoint with every derivative
// gradient
void case_2d_hermite()
{
    int const nc1 = 12, nc2 = 10, NB = 2;
    Lcg r{5};
    std::vector<double> br(nc1 + 1);
    br[0] = 0.0;
    for (int i = 1; i <= nc1; ++i) br[i] = br[i - 1] + 0.5 + r.next();
    for (auto& b : br) b /= br[nc1];
    auto const H = ddcx::SplineBuilderClosure::HERMITE;
    ddcx::NonUniformBSplines bs1(3, br, false);
    ddcx::UniformBSplines bs2(5, 0.0, 1.0, nc2, false);
    ddcx::SplineBuilder b1(bs1, H, H);
    ddcx::SplineBuilder b2(bs2, H, H);
    ddcx::SplineBuilder2D builder(b1, b2, bs1.size());
    int64_t const n1 = b1.npoints(), n2 = b2.npoints(), s1 = bs1.size(), s2 = bs2.size();
    int64_t const e1 = builder.nbe1(), e2 = builder.nbe2();
    auto rnd = [&](int64_t n) {
        std::vector<double> v(static_cast<size_t>(n));
        for (auto& x : v) x = r.next() - 0.5;
        return v;
    };
    std::vector<double> vals = rnd(NB * n1 * n2), gs = rnd(NB * s1 * s2);
    std::vector<std::vector<double>> d1 = {rnd(NB * e1 * n2), rnd(NB * e1 * n2)}, d2 = {rnd(NB * n1 * e2), rnd(NB * n1 * e2)};
    std::vector<std::vector<double>> m = {rnd(NB * e1 * e2), rnd(NB * e1 * e2), rnd(NB * e1 * e2), rnd(NB * e1 * e2)};
    DeviceArray dv(vals), dc(size_t(NB * s1 * s2)), dws(size_t(builder.workspace_size(NB))), dgs(gs);
    DeviceArray dd1a(d1[0]), dd1b(d1[1]), dd2a(d2[0]), dd2b(d2[1]), dma(m[0]), dmb(m[1]), dmc(m[2]), dmd(m[3]);
    builder(dc.p, dv.p, NB, dws.p, dd1a.p, dd1b.p, dd2a.p, dd2b.p, dma.p, dmb.p, dmc.p, dmd.p);
    DeviceArray gv(size_t(NB * n1 * n2)), g1a(size_t(NB * e1 * n2)), g1b(size_t(NB * e1 * n2));
    DeviceArray g2a(size_t(NB * n1 * e2)), g2b(size_t(NB * n1 * e2));
    DeviceArray gma(size_t(NB * e1 * e2)), gmb(size_t(NB * e1 * e2)), gmc(size_t(NB * e1 * e2)), gmd(size_t(NB * e1 * e2));
    builder.build_vjp(dgs.p, gv.p, NB, dws.p, g1a.p, g1b.p, g2a.p, g2b.p, gma.p, gmb.p, gmc.p, gmd.p);
    dump("2dh_breaks", br);
    dump("2dh_vals", vals);
    dump("2dh_gs", gs);
    char const* dn[] = {"2dh_dmin1", "2dh_dmax1", "2dh_dmin2", "2dh_dmax2", "2dh_mm", "2dh_Mm", "2dh_mM", "2dh_MM"};
    std::vector<double> const* dv8[] = {&d1[0], &d1[1], &d2[0], &d2[1], &m[0], &m[1], &m[2], &m[3]};
    for (int k = 0; k < 8; ++k) dump(dn[k], *dv8[k]);
    dump("2dh_coef", dc.get());
    DeviceArray const* g9[] = {&gv, &g1a, &g1b, &g2a, &g2b, &gma, &gmb, &gmc, &gmd};
    char const* gn[] = {"2dh_gvals", "2dh_gdmin1", "2dh_gdmax1", "2dh_gdmin2", "2dh_gdmax2",
                        "2dh_gmm", "2dh_gMm", "2dh_gmM", "2dh_gMM"};
    for (int k = 0; k < 9; ++k) dump(gn[k], g9[k]->get());
}

// 3D without derivative inputs: non-uniform cubic Greville (9 cells) x
// uniform periodic quintic (8 cells) x uniform quartic Greville (7 cells), 2
// batch entries; SplineBuilder3D through its two workspaces, and its adjoint
void case_3d()
{
    int const NB = 2;
    Lcg r{7};
    std::vector<double> br(10);
    br[0] = 0.0;
    for (int i = 1; i <= 9; ++i) br[i] = br[i - 1] + 0.5 + r.next();
    for (auto& b : br) b /= br[9];
    auto const G = ddcx::SplineBuilderClosure::GREVILLE, P = ddcx::SplineBuilderClosure::PERIODIC;
    ddcx::NonUniformBSplines bs1(3, br, false);
    ddcx::UniformBSplines bs2(5, 0.0, 1.0, 8, true);
    ddcx::UniformBSplines bs3(4, 0.0, 1.0, 7, false);
    ddcx::SplineBuilder b1(bs1, G, G), b2(bs2, P, P), b3(bs3, G, G);
    ddcx::SplineBuilder3D builder(b1, b2, b3, bs1.size(), bs2.size(), bs3.size());
    size_t const nv = size_t(NB) * b1.npoints() * b2.npoints() * b3.npoints();
    size_t const nc = size_t(NB) * bs1.size() * bs2.size() * bs3.size();
    std::vector<double> vals(nv), gs(nc);
    for (auto& x : vals) x = r.next() - 0.5;
    for (auto& x : gs) x = r.next() - 0.5;
    DeviceArray dv(vals), dc(nc), dgs(gs), dgv(nv);
    DeviceArray w1(static_cast<size_t>(builder.work1_size(NB))), w2(static_cast<size_t>(builder.work2_size(NB)));
    builder(dc.p, dv.p, NB, w1.p, w2.p);
    builder.build_vjp(dgs.p, dgv.p, NB, w1.p, w2.p);
    dump("3d_breaks", br);
    dump("3d_vals", vals);
    dump("3d_gs", gs);
    dump("3d_coef", dc.get());
    dump("3d_gvals", dgv.get());
}

// 3D with Hermite closures in every dimension: non-uniform cubic (9 cells) x
// uniform quintic (8 cells) x uniform quartic (7 cells), 2 batch entries;
// SplineBuilder3D with every derivative, edge and corner input
void case_3d_hermite()
{
    int const NB = 2;
    Lcg r{6};
    std::vector<double> br(10);
    br[0] = 0.0;
    for (int i = 1; i <= 9; ++i) br[i] = br[i - 1] + 0.5 + r.next();
    for (auto& b : br) b /= br[9];
    auto const H = ddcx::SplineBuilderClosure::HERMITE;
    ddcx::NonUniformBSplines bs1(3, br, false);
    ddcx::UniformBSplines bs2(5, 0.0, 1.0, 8, false);
    ddcx::UniformBSplines bs3(4, 0.0, 1.0, 7, false);
    ddcx::SplineBuilder b1(bs1, H, H), b2(bs2, H, H), b3(bs3, H, H);
    ddcx::SplineBuilder3D builder(b1, b2, b3, bs1.size(), bs2.size(), bs3.size());
    int64_t const n1 = b1.npoints(), n2 = b2.npoints(), n3 = b3.npoints();
    int64_t const e1 = std::max(b1.info().nbv_lower, b1.info().nbv_upper);
    int64_t const e2 = std::max(b2.info().nbv_lower, b2.info().nbv_upper);
    int64_t const e3 = std::max(b3.info().nbv_lower, b3.info().nbv_upper);
    std::vector<std::unique_ptr<DeviceArray>> keep;
    auto input = [&](std::string const& name, int64_t a, int64_t b, int64_t c) {
        std::vector<double> v(static_cast<size_t>(NB * a * b * c));
        for (auto& x : v) x = r.next() - 0.5;
        dump("3dh_" + name, v);
        keep.push_back(std::make_unique<DeviceArray>(v));
        return static_cast<const double*>(keep.back()->p);
    };
    char const* sd[2] = {"min", "max"};
    const double* vals = input("vals", n1, n2, n3);
    ddcx::SplineBuilder3DDerivs d;
    d.derivs_min1 = input("derivs_min1", e1, n2, n3);
    d.derivs_max1 = input("derivs_max1", e1, n2, n3);
    d.derivs_min2 = input("derivs_min2", n1, e2, n3);
    d.derivs_max2 = input("derivs_max2", n1, e2, n3);
    d.derivs_min3 = input("derivs_min3", n1, n2, e3);
    d.derivs_max3 = input("derivs_max3", n1, n2, e3);
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            d.mixed12[a][b] = input(std::string("mixed_derivs_") + sd[a] + "1_" + sd[b] + "2", e1, e2, n3);
            d.mixed23[a][b] = input(std::string("mixed_derivs_") + sd[a] + "2_" + sd[b] + "3", n1, e2, e3);
            d.mixed13[a][b] = input(std::string("mixed_derivs_") + sd[a] + "1_" + sd[b] + "3", e1, n2, e3);
            for (int c = 0; c < 2; ++c)
                d.mixed123[a][b][c] =
                    input(std::string("mixed_derivs_") + sd[a] + "1_" + sd[b] + "2_" + sd[c] + "3", e1, e2, e3);
        }
    DeviceArray coef(static_cast<size_t>(NB * bs1.size() * bs2.size() * bs3.size()));
    DeviceArray work(static_cast<size_t>(builder.hermite_workspace_size(NB)));
    builder(coef.p, vals, NB, work.p, d);
    dump("3dh_breaks", br);
    dump("3dh_coef", coef.get());
    // the adjoint: a random coefficient gradient, every input's gradient
    std::vector<double> gs(static_cast<size_t>(NB * bs1.size() * bs2.size() * bs3.size()));
    for (auto& x : gs) x = r.next() - 0.5;
    dump("3dh_gs", gs);
    DeviceArray dgs(gs);
    std::vector<std::pair<std::string, std::unique_ptr<DeviceArray>>> grads;
    auto grad = [&](std::string const& name, int64_t a, int64_t b, int64_t c) {
        grads.emplace_back(name, std::make_unique<DeviceArray>(static_cast<size_t>(NB * a * b * c)));
        return grads.back().second->p;
    };
    double* gvals = grad("vals", n1, n2, n3);
    ddcx::SplineBuilder3DGrads g;
    g.derivs_min1 = grad("derivs_min1", e1, n2, n3);
    g.derivs_max1 = grad("derivs_max1", e1, n2, n3);
    g.derivs_min2 = grad("derivs_min2", n1, e2, n3);
    g.derivs_max2 = grad("derivs_max2", n1, e2, n3);
    g.derivs_min3 = grad("derivs_min3", n1, n2, e3);
    g.derivs_max3 = grad("derivs_max3", n1, n2, e3);
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            g.mixed12[a][b] = grad(std::string("mixed_derivs_") + sd[a] + "1_" + sd[b] + "2", e1, e2, n3);
            g.mixed23[a][b] = grad(std::string("mixed_derivs_") + sd[a] + "2_" + sd[b] + "3", n1, e2, e3);
            g.mixed13[a][b] = grad(std::string("mixed_derivs_") + sd[a] + "1_" + sd[b] + "3", e1, n2, e3);
            for (int c = 0; c < 2; ++c)
                g.mixed123[a][b][c] =
                    grad(std::string("mixed_derivs_") + sd[a] + "1_" + sd[b] + "2_" + sd[c] + "3", e1, e2, e3);
        }
    builder.build_vjp(dgs.p, gvals, NB, work.p, g);
    for (auto const& [name, arr] : grads) dump("3dh_g" + name, arr->get());
}

} // namespace

int main(int argc, char** argv)
{
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s OUTPUT_DIR\n", argv[0]);
        return 2;
    }
    g_dir = argv[1];
    try {
        std::printf("facade_device: 1D batch-contiguous\n");
        std::fflush(stdout);
        case_1d(true);
        std::printf("facade_device: 1D interpolation-contiguous\n");
        std::fflush(stdout);
        case_1d(false);
        std::printf("facade_device: constant rules with eval_pos\n");
        std::fflush(stdout);
        case_rules();
        std::printf("facade_device: 2D\n");
        std::fflush(stdout);
        case_2d();
        std::printf("facade_device: 2D Hermite\n");
        std::fflush(stdout);
        case_2d_hermite();
        std::printf("facade_device: 3D\n");
        std::fflush(stdout);
        case_3d();
        std::printf("facade_device: 3D Hermite\n");
        std::fflush(stdout);
        case_3d_hermite();
    } catch (std::exception const& e) {
        std::fprintf(stderr, "facade_device: %s\n", e.what());
        return 1;
    }
    std::printf("facade_device: done\n");
    return 0;
}
