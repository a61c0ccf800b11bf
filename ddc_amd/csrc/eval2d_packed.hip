// Batched 2D evaluation (SplineEvaluator2D::operator() / deriv*,
// spline_evaluator_2d.hpp:370-825) for many points per batch entry, through a
// packed copy of the coefficients.
//
// Scattered points gather a (d1+1) x (d2+1) coefficient patch each.  From the
// plain (size1, size2) array that is d1+1 separate rows, each its own cache
// line, and with 2^24 points per batch entry the 8 MiB array does not stay in
// one XCD's L2.  So the call first packs, for every cell (i1, i2), the patch
// that cell's points use into one contiguous, 16-byte aligned record of
// P = (d1+1)(d2+1) doubles (padded to even), in a caller workspace; then
//   1. each lane of a wave takes one point: rules, cell indices and basis
//      values (the same routines as eval2d_kernel, with the knots and cell
//      brackets of a non-uniform dimension staged in LDS), parked in LDS;
//   2. groups of 8 lanes take one point each and load its record as 16-byte
//      pairs (one cache line per point for cubic x cubic); the records of a
//      wave's 64 points are all requested before the first is used, so each
//      wave keeps 8 cache-line gathers in flight;
//   3. each lane multiplies its record pairs by the basis products, one
//      partial sum per point of its group, and three DPP exchanges within
//      the group (mirrored lane, then quad permutations) fold the 8 x 8
//      partial sums so that lane k ends with point k's value;
//   4. each lane stores its point's value (coalesced).
// The terms are summed in a fixed order, so results do not depend on
// scheduling; they differ from eval2d_kernel's by rounding only.
#include <algorithm>

#include "internal.hpp"

namespace ddcx {

namespace {

constexpr int kPackThreads = 256;
constexpr int kEvalThreads = 256;
constexpr int kGroup = 8; // lanes per point in the gather
constexpr size_t kTabLds = 64 * 1024; // LDS for staged knots and brackets

__host__ __device__ constexpr int patch_len(int d1, int d2)
{
    return ((d1 + 1) * (d2 + 1) + 1) / 2 * 2;
}

// One thread per (batch entry, cell, pair of record elements): the record of
// cell (i1, i2) lists its patch row by row, element k being coefficient
// (i1 + k / (d2+1), i2 + k % (d2+1)), then the zero pad; consecutive threads
// write consecutive 16-byte pairs.
template <int D1, int D2>
__global__ void __launch_bounds__(kPackThreads) pack2d_kernel(const double* __restrict__ coef, int64_t cbs,
                                                              int64_t st1, int64_t st2, int nc1, int nc2,
                                                              int64_t nbatch, double* __restrict__ packed)
{
    constexpr int NPAIR = patch_len(D1, D2) / 2;
    constexpr int NE = (D1 + 1) * (D2 + 1);
    int64_t const t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    int64_t const ncell = int64_t(nc1) * nc2;
    if (t >= ncell * nbatch * NPAIR) return;
    int64_t const cell = t / NPAIR; // (l * nc1 + i1) * nc2 + i2
    int const e = 2 * int(t - cell * NPAIR);
    int64_t const l = cell / ncell;
    int64_t const r = cell - l * ncell;
    int64_t const i1 = r / nc2, i2 = r - i1 * nc2;
    const double* c = coef + l * cbs;
    auto elem = [&](int k) {
        if (k >= NE) return 0.0; // the pad element
        return c[(i1 + k / (D2 + 1)) * st1 + (i2 + k % (D2 + 1)) * st2];
    };
    reinterpret_cast<double2*>(packed)[t] = make_double2(elem(e), elem(e + 1));
}

// v of the lane that DPP control CTRL names (quad permutations, row mirrors)
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    int const lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xf, 0xf, true);
    int const hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}

// One step of the transposed sum over a group of 8 lanes: each lane pairs
// with the lane that DPP control CTRL names; the lane with `upper` false
// keeps the first half of its N partial sums and adds its partner's first
// half, the other keeps and adds the second half.
template <int N, int CTRL>
__device__ __forceinline__ void fold_half(double (&v)[8], bool upper)
{
#pragma unroll
    for (int i = 0; i < N / 2; ++i) {
        double const send = upper ? v[i] : v[i + N / 2];
        double const keep = upper ? v[i + N / 2] : v[i];
        v[i] = keep + dpp_f64<CTRL>(send);
    }
}

// Persistent: workgroup g stages the knot vector and bucket brackets of each
// non-uniform dimension in LDS, then takes tiles g, g + G, ... of kEvalThreads points, tile
// t being points (t % tiles) * kEvalThreads ... of batch entry t / tiles.
template <int D1, int D2, int O1, int O2>
__global__ void __launch_bounds__(kEvalThreads) eval2d_packed_kernel(BSplinesView b1, BSplinesView b2, Rules2D rules,
                                                                     double* __restrict__ out,
                                                                     const double* __restrict__ coords, int64_t npts,
                                                                     const double* __restrict__ packed, int nc1,
                                                                     int nc2, int64_t nbatch)
{
    constexpr int P = patch_len(D1, D2);
    constexpr int NPAIR = P / 2;
    constexpr int NB = D1 + 1 + D2 + 1;
    constexpr int WAVES = kEvalThreads / 64;
    __shared__ double sb[WAVES][64][NB];  // basis values of dimension 1, then 2
    __shared__ int64_t srec[WAVES][64];   // record index
    __shared__ double so[WAVES][64];
    extern __shared__ double lds_tabs[]; // staged knots and cell brackets
    stage_tables_2d(b1, b2, lds_tabs, threadIdx.x, kEvalThreads);
    __syncthreads();
    int const wv = threadIdx.x / 64, lane = threadIdx.x % 64;
    int const sub = lane % kGroup, grp = lane / kGroup;
    int64_t const tiles = (npts + kEvalThreads - 1) / kEvalThreads;
    // tile t = l * tiles + tt of this workgroup, advanced by gridDim.x
    int64_t l = blockIdx.x / tiles, tt = blockIdx.x - (blockIdx.x / tiles) * tiles;
    for (; l < nbatch; tt += gridDim.x) {
        while (tt >= tiles && l < nbatch) {
            tt -= tiles;
            ++l;
        }
        if (l >= nbatch) break;
        int64_t const p0 = tt * kEvalThreads + wv * 64;
        if (p0 >= npts) continue; // wave-uniform
        bool keep = true;
        {
            int64_t const p = p0 + lane < npts ? p0 + lane : npts - 1;
            double2 const xy = reinterpret_cast<const double2*>(coords)[l * npts + p];
            double x1 = xy.x, x2 = xy.y;
            if constexpr (O1 == 0 && O2 == 0) {
                bool in1, in2;
                keep = place2d(b1, b2, rules, x1, x2, in1, in2);
            }
            double v1[D1 + 1], v2[D2 + 1];
            int j1, j2;
            if constexpr (O1 == 0) j1 = eval_basis<D1, true>(b1, b1.knots, x1, v1);
            else {
                double der[(D1 + 1) * (O1 + 1)];
                j1 = eval_basis_and_n_derivs<D1, O1, true>(b1, b1.knots, x1, der);
#pragma unroll
                for (int i = 0; i < D1 + 1; ++i) v1[i] = der[i * (O1 + 1) + O1];
            }
            if constexpr (O2 == 0) j2 = eval_basis<D2, true>(b2, b2.knots, x2, v2);
            else {
                double der[(D2 + 1) * (O2 + 1)];
                j2 = eval_basis_and_n_derivs<D2, O2, true>(b2, b2.knots, x2, der);
#pragma unroll
                for (int i = 0; i < D2 + 1; ++i) v2[i] = der[i * (O2 + 1) + O2];
            }
#pragma unroll
            for (int i = 0; i < D1 + 1; ++i) sb[wv][lane][i] = v1[i];
#pragma unroll
            for (int j = 0; j < D2 + 1; ++j) sb[wv][lane][D1 + 1 + j] = v2[j];
            // a point the Null rule sends to 0 reads record 0 and stores 0
            srec[wv][lane] = keep ? (l * nc1 + j1) * int64_t(nc2) + j2 : 0;
        }
        __builtin_amdgcn_wave_barrier();
        // lane (grp, sub) forms, for each of the 8 points it = 0..7 of its
        // group (point it * 8 + grp of the slice), the partial sum over its
        // record elements; the records of BATCH points are loaded before any
        // is used
        constexpr int KL = (NPAIR + kGroup - 1) / kGroup; // 16-byte loads per lane and point
        constexpr int BATCH = KL == 1 ? 8 : (KL == 2 ? 4 : 2);
        double part[8];
#pragma unroll
        for (int i0 = 0; i0 < 64 / kGroup; i0 += BATCH) {
            double2 c[BATCH][KL];
#pragma unroll
            for (int u = 0; u < BATCH; ++u) {
                int const q = (i0 + u) * kGroup + grp;
                const double2* r2 = reinterpret_cast<const double2*>(packed + srec[wv][q] * P);
#pragma unroll
                for (int k = 0; k < KL; ++k) {
                    int const e2 = sub + k * kGroup;
                    c[u][k] = e2 < NPAIR ? r2[e2] : make_double2(0.0, 0.0);
                }
            }
#pragma unroll
            for (int u = 0; u < BATCH; ++u) {
                int const q = (i0 + u) * kGroup + grp;
                double pt = 0.0;
#pragma unroll
                for (int k = 0; k < KL; ++k) {
                    int const e2 = sub + k * kGroup;
                    if (e2 >= NPAIR) break;
                    int const e = 2 * e2;
                    // element e is (row e / (D2+1), column e % (D2+1)); the pad element has a zero coefficient
                    int const ia = e / (D2 + 1), ja = e % (D2 + 1);
                    int const ib = (e + 1) / (D2 + 1), jb = (e + 1) % (D2 + 1);
                    double const wa = sb[wv][q][ia] * sb[wv][q][D1 + 1 + ja];
                    double const wb = (e + 1 < (D1 + 1) * (D2 + 1)) ? sb[wv][q][ib] * sb[wv][q][D1 + 1 + jb] : 0.0;
                    pt = fma(c[u][k].x, wa, pt);
                    pt = fma(c[u][k].y, wb, pt);
                }
                part[i0 + u] = pt;
            }
        }
        // transposed sum over the group: lane sub ends with the sum of point
        // it = sub (partners: mirrored lane 7 - sub, then sub ^ 2, then sub ^ 1)
        fold_half<8, 0x141>(part, sub >= 4);   // row_half_mirror
        fold_half<4, 0x4E>(part, (sub & 2) != 0); // quad_perm [2, 3, 0, 1]
        fold_half<2, 0xB1>(part, (sub & 1) != 0); // quad_perm [1, 0, 3, 2]
        so[wv][sub * kGroup + grp] = part[0];
        __builtin_amdgcn_wave_barrier();
        if (p0 + lane < npts) out[l * npts + p0 + lane] = keep ? so[wv][lane] : 0.0;
        // the next tile's LDS writes come after every lane's read of this one
        __builtin_amdgcn_wave_barrier();
    }
}

template <class F>
bool dispatch_2d(int d1, int d2, int o1, int o2, F&& f)
{
    bool done = false;
    auto go = [&]<int D1, int D2>() {
        auto sel = [&]<int O1, int O2>() {
            if constexpr (O1 <= D1 && O2 <= D2) {
                if (!done && o1 == O1 && o2 == O2) {
                    f.template operator()<D1, D2, O1, O2>();
                    done = true;
                }
            }
        };
        sel.template operator()<0, 0>();
        sel.template operator()<1, 0>();
        sel.template operator()<0, 1>();
        sel.template operator()<1, 1>();
        sel.template operator()<2, 0>();
        sel.template operator()<0, 2>();
        sel.template operator()<2, 1>();
        sel.template operator()<1, 2>();
        sel.template operator()<2, 2>();
    };
    auto outer = [&]<int D1>() {
        switch (d2) {
        case 1: go.template operator()<D1, 1>(); break;
        case 2: go.template operator()<D1, 2>(); break;
        case 3: go.template operator()<D1, 3>(); break;
        case 4: go.template operator()<D1, 4>(); break;
        case 5: go.template operator()<D1, 5>(); break;
        default: break;
        }
    };
    switch (d1) {
    case 1: outer.template operator()<1>(); break;
    case 2: outer.template operator()<2>(); break;
    case 3: outer.template operator()<3>(); break;
    case 4: outer.template operator()<4>(); break;
    case 5: outer.template operator()<5>(); break;
    default: break;
    }
    return done;
}

} // namespace

int64_t eval2d_packed_record_bytes(BSplinesPlan const& bp1, BSplinesPlan const& bp2)
{
    int const d1 = bp1.view.degree, d2 = bp2.view.degree;
    return int64_t(bp1.view.ncells) * bp2.view.ncells * patch_len(d1, d2) * int64_t(sizeof(double));
}

void launch_eval2d_packed(BSplinesPlan const& bp1, BSplinesPlan const& bp2, Rules2D const& rules, int o1, int o2,
                          double* out, const double* coords, int64_t npts, const double* coef, int64_t cbs,
                          int64_t st1, int64_t st2, int64_t nbatch, void* work, int64_t work_bytes,
                          hipStream_t stream)
{
    if (npts * nbatch == 0) return;
    bp1.ensure_device();
    bp2.ensure_device();
    BSplinesView const b1 = bp1.device_view(), b2 = bp2.device_view();
    int const nc1 = b1.ncells, nc2 = b2.ncells;
    int64_t const rec_bytes = eval2d_packed_record_bytes(bp1, bp2);
    if (reinterpret_cast<uintptr_t>(work) % 16 != 0)
        throw Error(DDCX_ERR_INVALID, "packed 2D evaluation: workspace must be 16-byte aligned");
    int64_t const chunk = work_bytes / rec_bytes; // batch entries packed at a time
    if (chunk < 1) throw Error(DDCX_ERR_INVALID, "packed 2D evaluation: workspace smaller than one batch entry");
    double* packed = static_cast<double*>(work);
    // the knots and brackets of the non-uniform dimensions go to LDS; grids
    // whose tables do not fit take the direct kernel
    if (!staged_tables_fit_2d(b1, b2, kTabLds)) {
        launch_eval2d(bp1, bp2, rules, o1, o2, out, coords, npts, coef, cbs, st1, st2, nbatch, stream);
        return;
    }
    size_t const lds = staged_tables_bytes_2d(b1, b2);
    int64_t const tiles = (npts + kEvalThreads - 1) / kEvalThreads;
    bool const ok = dispatch_2d(b1.degree, b2.degree, o1, o2, [&]<int D1, int D2, int O1, int O2>() {
        auto kern = eval2d_packed_kernel<D1, D2, O1, O2>;
        if (lds > 0)
            hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)),
                      "hipFuncSetAttribute");
        int per_cu = 0;
        hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kEvalThreads, lds),
                  "hipOccupancyMaxActiveBlocksPerMultiprocessor");
        int64_t const resident = int64_t(std::max(per_cu, 1)) * cu_count();
        for (int64_t l0 = 0; l0 < nbatch; l0 += chunk) {
            int64_t const nb = std::min(chunk, nbatch - l0);
            int64_t const nthreads = int64_t(nc1) * nc2 * (patch_len(D1, D2) / 2) * nb;
            hipLaunchKernelGGL((pack2d_kernel<D1, D2>), dim3(unsigned((nthreads + kPackThreads - 1) / kPackThreads)),
                               dim3(kPackThreads), 0, stream, coef + l0 * cbs, cbs, st1, st2, nc1, nc2, nb, packed);
            unsigned const grid = unsigned(std::min<int64_t>(tiles * nb, resident));
            hipLaunchKernelGGL(kern, dim3(grid), dim3(kEvalThreads), lds, stream, b1, b2, rules, out + l0 * npts,
                               coords + 2 * l0 * npts, npts, packed, nc1, nc2, nb);
        }
    });
    if (!ok) throw Error(DDCX_ERR_UNSUPPORTED, "2D evaluation supports degrees 1..5 and derivative orders <= 2");
    hip_check(hipGetLastError(), "packed 2D evaluation launch");
}

} // namespace ddcx
