// LDS-staged batched evaluation (SplineEvaluator::operator() / deriv(),
// spline_evaluator.hpp:263-478).
//
// Random evaluation points gather the coefficients of their line at random,
// so a workgroup takes LB consecutive lines, stages their coefficient tile
// (size() x LB) and, for non-uniform bases, the knot vector and the cell-search
// brackets in LDS, and then evaluates all the points of those lines with
// coalesced coordinate loads and output stores.  HBM sees every coefficient,
// coordinate and output once.  Two thread mappings:
//  - batch-contiguous (LINE_MAJOR = false): lanes take consecutive lines, so
//    tile rows and point rows are read as LB*8-byte runs;
//  - interpolation-contiguous (LINE_MAJOR = true): the tile is read line by
//    line (size() contiguous doubles each) and the NT/LB threads of a line
//    take consecutive points of it.
// Values on plain batch-contiguous arrays take eval_stream_kernel, a
// persistent form that loads the next tile while the current one is
// evaluated.
#include "internal.hpp"

namespace ddcx {

namespace {

__device__ __forceinline__ int64_t lay(Layout const& L, int64_t r, int64_t l)
{
    return r * L.rs + (l % L.inner) * L.is + (l / L.inner) * L.os;
}

template <int D, int ORDER, int KIND>
__device__ __forceinline__ double point_value(BSplinesView const& b, const double* kn, Rules1D rules, double x,
                                              const double* tile, int pitch)
{
    double v[D + 1];
    int jmin;
    if constexpr (ORDER == 0) {
        if (b.periodic) {
            if (x < b.rmin || x > b.rmax) x -= floor((x - b.rmin) / b.length) * b.length;
        } else if (x < b.rmin) {
            if (rules.lower == DDCX_EXTRAP_NULL) return 0.0;
            x = rules.pos_lower;
        } else if (x > b.rmax) {
            if (rules.upper == DDCX_EXTRAP_NULL) return 0.0;
            x = rules.pos_upper;
        }
        jmin = eval_basis<D, true, KIND>(b, kn, x, v);
    } else {
        double der[(D + 1) * (ORDER + 1)];
        jmin = eval_basis_and_n_derivs<D, ORDER, true, KIND>(b, kn, x, der);
#pragma unroll
        for (int i = 0; i < D + 1; ++i) v[i] = der[i * (ORDER + 1) + ORDER];
    }
    double y = 0.0;
#pragma unroll
    for (int i = 0; i < D + 1; ++i) y += tile[(jmin + i) * pitch] * v[i];
    return y;
}

template <int D, int ORDER, int LB, int NT, bool LINE_MAJOR, int KIND>
__global__ void __launch_bounds__(NT) eval_tiled_kernel(BSplinesView b, Rules1D rules, double* __restrict__ out,
                                                        const double* __restrict__ coords, Layout xl, Layout pl,
                                                        int64_t npts,
                                                        const double* __restrict__ coef, Layout cl, int64_t nbatch,
                                                        int ncoef, int pitch, int nknots)
{
    extern __shared__ double smem[];
    double* tile = smem;
    double* kn = smem + size_t(ncoef) * pitch;
    uint32_t* cells = reinterpret_cast<uint32_t*>(kn + nknots);
    constexpr int ROWS = NT / LB;
    // points per pipeline stage per thread: as many as the registers hold
    // without spilling (derivative tables of non-uniform bases are the largest)
    constexpr int U = ORDER == 0 ? 8
                      : KIND == kUniformBasis ? (D <= 4 ? 8 : 4)
                                              : (D <= 1 ? 8 : (D <= 3 ? 4 : 2));
    constexpr int TC = 20;         // tile rows per thread per load chunk (one chunk up to 20 * ROWS rows)
    int const tid = threadIdx.x;
    int const ln = LINE_MAJOR ? tid / ROWS : tid % LB;
    int const prow = LINE_MAJOR ? tid % ROWS : tid / LB;
    int64_t const line0 = int64_t(blockIdx.x) * LB;
    int64_t const line = line0 + ln;
    bool const active = line < nbatch;
    int64_t const lc = active ? line : 0;
    int64_t const cb = lay(cl, 0, lc);
    int64_t const pb = lay(pl, 0, lc);
    // Point i of the line is at cp[i * prs].  Indices past the end are clamped
    // to the last point: such a slot loads the last coordinate and stores the
    // last point's own value again, so every load and store of the loop below
    // is unconditional and the loads of the next stage stay in flight while
    // the current stage is evaluated.
    int const last = int(npts) - 1;
    uint32_t const prs = uint32_t(pl.rs), xrs = uint32_t(xl.rs);
    const double* __restrict__ cp = coords + lay(xl, 0, lc);
    double* __restrict__ op = out + pb;
    auto at = [&](int ii) { return uint64_t(uint32_t(min(ii, last))) * prs; };
    auto atx = [&](int ii) { return uint64_t(uint32_t(min(ii, last))) * xrs; };
    // first batch of coordinates, then the whole coefficient tile: every load
    // of the tile is in flight before the first LDS write
    double xa[U];
#pragma unroll
    for (int u = 0; u < U; ++u) xa[u] = __builtin_nontemporal_load(&cp[atx(prow + u * ROWS)]);
    // the tile is loaded in chunks of TC elements per thread, every load of a
    // chunk in flight before its LDS writes
    if constexpr (LINE_MAJOR) {
        int const nl = int(min<int64_t>(LB, nbatch - line0));
        int const nel = nl * ncoef;
        for (int e0 = tid; e0 < nel; e0 += TC * NT) {
            double tv[TC];
#pragma unroll
            for (int k = 0; k < TC; ++k) {
                int const e = e0 + k * NT;
                int const l = e / ncoef, r = e - l * ncoef;
                tv[k] = e < nel ? coef[lay(cl, r, line0 + l)] : 0.0;
            }
#pragma unroll
            for (int k = 0; k < TC; ++k) {
                int const e = e0 + k * NT;
                int const l = e / ncoef, r = e - l * ncoef;
                if (e < nel) tile[r * pitch + l] = tv[k];
            }
        }
    } else {
        for (int c0 = prow; c0 < ncoef; c0 += TC * ROWS) {
            double tv[TC];
#pragma unroll
            for (int k = 0; k < TC; ++k) {
                int const r = c0 + k * ROWS;
                tv[k] = r < ncoef ? coef[cb + int64_t(r) * cl.rs] : 0.0;
            }
#pragma unroll
            for (int k = 0; k < TC; ++k) {
                int const r = c0 + k * ROWS;
                if (r < ncoef) tile[r * pitch + ln] = tv[k];
            }
        }
    }
    if (!is_uniform_kind(KIND, b.uniform)) {
        for (int k = tid; k < nknots; k += NT) kn[k] = b.knots[k];
        for (int k = tid; k < b.nbucket; k += NT) cells[k] = b.cells[k];
    }
    __syncthreads();
    if (!active) return;
    const double* knots = is_uniform_kind(KIND, b.uniform) ? nullptr : kn;
    BSplinesView bl = b;
    bl.cells = b.nbucket > 0 ? cells : nullptr;
    const double* tl = tile + ln;
    // software pipeline: the next batch of coordinates is in flight while the
    // current one is evaluated and stored
    for (int i = prow; i <= last; i += U * ROWS) {
        double xb[U];
#pragma unroll
        for (int u = 0; u < U; ++u) xb[u] = __builtin_nontemporal_load(&cp[atx(i + (U + u) * ROWS)]);
        double y[U];
#pragma unroll
        for (int u = 0; u < U; ++u) y[u] = point_value<D, ORDER, KIND>(bl, knots, rules, xa[u], tl, pitch);
#pragma unroll
        for (int u = 0; u < U; ++u) __builtin_nontemporal_store(y[u], &op[at(i + u * ROWS)]);
#pragma unroll
        for (int u = 0; u < U; ++u) xa[u] = xb[u];
    }
}

// Buffer resources (raw buffers, stride 0) for the streaming kernel: a load
// at a byte offset at or past num_records returns 0 and a store there is
// dropped, which is how the kernel handles the points past the end of a line.
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
constexpr int kBufferWord3 = 0x00020000; // 32-bit data format, gfx9 family

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_of(const double* p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(p), 0, int(bytes), kBufferWord3);
}

// Cache-policy bit NT (non-temporal) of the buffer instructions: every
// coordinate, coefficient and output of the streaming kernel is touched once,
// so its lines need not displace what the next kernel reads.
constexpr int kNonTemporal = 2;

template <int AUX = 0>
__device__ __forceinline__ double buffer_load(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, int(off), 0, AUX));
}

template <int AUX = 0>
__device__ __forceinline__ void buffer_store(double v, __amdgpu_buffer_rsrc_t r, uint32_t off)
{
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, int(off), 0, AUX);
}

// Persistent form of eval_tiled_kernel for values with plain batch-contiguous
// arrays (element (row, line) at row * rs + line, 16-line tiles, NT threads:
// 512 for uniform bases, 768 for non-uniform ones).  Each workgroup walks the tiles blockIdx.x, blockIdx.x +
// gridDim.x, ...: while the points of one tile are evaluated, the coefficients
// of its next tile are loaded into registers (PR rows per thread, one slice per
// stage of U points), and they go to LDS once the current tile has been read.
// Coefficient loads, point loads and stores therefore overlap inside every CU
// instead of alternating in lockstep across the chip.  Every load and store of
// a tile is unconditional, so the compiler keeps each stage's loads in flight
// with exact vmcnt waits: points past the end of a line fall outside the
// buffers (loads give 0, stores are dropped), rows past the end of the tile
// repeat the last row, and lines past the batch repeat the last line.
// Addresses are 32-bit byte offsets into buffers of at most 4 GiB.
template <int D, int ORDER, int KIND, int NT, int U, int NST, int PR>
__global__ void __launch_bounds__(NT) eval_stream_kernel(BSplinesView b, Rules1D rules,
                                                           double* __restrict__ out,
                                                           const double* __restrict__ coords, uint32_t prs,
                                                           uint32_t pbytes, const double* __restrict__ coef,
                                                           uint32_t crs, uint32_t cbytes, int64_t nbatch, int ncoef,
                                                           int pitch, int nknots, int64_t ntiles, int npts)
{
    constexpr int LB = 16, ROWS = NT / LB;
    constexpr int PS = (PR + NST - 1) / NST; // prefetch rows per stage
    extern __shared__ double smem[];
    double* tile = smem;
    double* kn = smem + size_t(ncoef) * pitch;
    uint32_t* cells = reinterpret_cast<uint32_t*>(kn + nknots);
    int const tid = threadIdx.x;
    int const ln = tid % LB;
    int const prow = tid / LB;
    uint32_t const lastr = uint32_t(ncoef - 1);
    __amdgpu_buffer_rsrc_t const rc = buffer_of(coords, pbytes);
    __amdgpu_buffer_rsrc_t const ro = buffer_of(out, pbytes);
    __amdgpu_buffer_rsrc_t const rk = buffer_of(coef, cbytes);
    // byte offsets of point slot p and coefficient row j of line l.  Slot p
    // of tile t is point (p + rot(t)) mod npts, slots past npts stay outside
    // the buffer; the rotation of 64 points per tile spreads the rows that the
    // workgroups read at the same time over more DRAM banks.
    int rot = 0;
    auto pt = [&](int p, uint32_t l) {
        int q = p + rot;
        q = q >= npts ? q - npts : q;
        q = p < npts ? q : p;
        return (uint32_t(q) * prs + l) * 8u;
    };
    auto cr = [&](int j, uint32_t l) { return (min(uint32_t(prow + j * ROWS), lastr) * crs + l) * 8u; };
    auto line_of = [&](int64_t t) { return uint32_t(min(t * LB + ln, nbatch - 1)); };

    if (!is_uniform_kind(KIND, b.uniform)) {
        for (int k = tid; k < nknots; k += NT) kn[k] = b.knots[k];
        for (int k = tid; k < b.nbucket; k += NT) cells[k] = b.cells[k];
    }
    const double* knots = is_uniform_kind(KIND, b.uniform) ? nullptr : kn;
    BSplinesView bl = b;
    bl.cells = b.nbucket > 0 ? cells : nullptr;

    int64_t const G = gridDim.x;
    int64_t t = blockIdx.x;
    uint32_t lt = line_of(t);
    auto rot_of = [&](int64_t tt) { return int((tt * 64) % npts); };
    rot = rot_of(t);
    double xa[U];
#pragma unroll
    for (int u = 0; u < U; ++u) xa[u] = buffer_load<kNonTemporal>(rc, pt(prow + u * ROWS, lt));
    {
        double tv[PR];
#pragma unroll
        for (int j = 0; j < PR; ++j) tv[j] = buffer_load<kNonTemporal>(rk, cr(j, lt));
        int const col = int(int64_t(lt) - t * LB);
#pragma unroll
        for (int j = 0; j < PR; ++j) {
            int const r = prow + j * ROWS;
            if (r < ncoef) tile[r * pitch + col] = tv[j];
        }
    }
    while (true) {
        int64_t const tn = t + G;
        bool const more = tn < ntiles;
        uint32_t const ltn = line_of(more ? tn : t);
        // after the last tile the prefetch reads empty buffers: every access
        // is out of range, so it costs no memory traffic
        __amdgpu_buffer_rsrc_t const rcn = buffer_of(coords, more ? pbytes : 0u);
        __amdgpu_buffer_rsrc_t const rkn = buffer_of(coef, more ? cbytes : 0u);
        const double* tl = tile + (int64_t(lt) - t * LB);
        __syncthreads(); // tile t is in LDS
        double pf[PR];
#pragma unroll
        for (int s = 0; s < NST; ++s) {
            double xb[U];
            if (s + 1 < NST) {
#pragma unroll
                for (int u = 0; u < U; ++u) xb[u] = buffer_load<kNonTemporal>(rc, pt((s + 1) * U * ROWS + prow + u * ROWS, lt));
            } else {
                int const r0 = rot;
                rot = rot_of(more ? tn : t);
#pragma unroll
                for (int u = 0; u < U; ++u) xb[u] = buffer_load<kNonTemporal>(rcn, pt(prow + u * ROWS, ltn));
                rot = r0;
            }
#pragma unroll
            for (int j = s * PS; j < (s + 1) * PS && j < PR; ++j) pf[j] = buffer_load<kNonTemporal>(rkn, cr(j, ltn));
            double y[U];
#pragma unroll
            for (int u = 0; u < U; ++u) y[u] = point_value<D, ORDER, KIND>(bl, knots, rules, xa[u], tl, pitch);
#pragma unroll
            for (int u = 0; u < U; ++u) buffer_store<kNonTemporal>(y[u], ro, pt(s * U * ROWS + prow + u * ROWS, lt));
#pragma unroll
            for (int u = 0; u < U; ++u) xa[u] = xb[u];
        }
        __syncthreads(); // every read of tile t is done
        if (!more) break;
        int const col = int(int64_t(ltn) - tn * LB);
#pragma unroll
        for (int j = 0; j < PR; ++j) {
            int const r = prow + j * ROWS;
            if (r < ncoef) tile[r * pitch + col] = pf[j];
        }
        t = tn;
        lt = ltn;
        rot = rot_of(t);
    }
}

// Adjoint of the value form on the same tiles: gx = w * f'(x) where the
// spline itself is evaluated (0 where an extrapolation rule applies), and the
// coefficient gradient of the LB lines accumulated in an LDS tile with LDS
// atomics, then written once (every row, so no separate zeroing).
template <int D, int LB, int NT, bool LINE_MAJOR>
__global__ void __launch_bounds__(NT) eval_vjp_tiled_kernel(BSplinesView b, Rules1D rules,
                                                            const double* __restrict__ coords,
                                                            const double* __restrict__ w, double* __restrict__ gx,
                                                            Layout pl, int64_t npts, const double* __restrict__ coef,
                                                            double* __restrict__ gcoef, Layout cl, int64_t nbatch,
                                                            int ncoef, int pitch, int nknots)
{
    extern __shared__ double smem[];
    double* tile = smem;
    double* gt = smem + size_t(ncoef) * pitch;
    double* kn = gt + size_t(ncoef) * pitch;
    uint32_t* cells = reinterpret_cast<uint32_t*>(kn + nknots);
    constexpr int ROWS = NT / LB;
    constexpr int U = 4;
    constexpr int TC = 16;
    int const tid = threadIdx.x;
    int const ln = LINE_MAJOR ? tid / ROWS : tid % LB;
    int const prow = LINE_MAJOR ? tid % ROWS : tid / LB;
    int64_t const line0 = int64_t(blockIdx.x) * LB;
    int64_t const line = line0 + ln;
    bool const active = line < nbatch;
    int64_t const lc = active ? line : 0;
    int64_t const cb = lay(cl, 0, lc);
    int64_t const pb = lay(pl, 0, lc);
    int const nl = int(min<int64_t>(LB, nbatch - line0));
    int const nel = nl * ncoef;
    // coefficient tile (as in eval_tiled_kernel) and a zeroed gradient tile
    if constexpr (LINE_MAJOR) {
        for (int e0 = tid; e0 < nel; e0 += TC * NT) {
            double tv[TC];
#pragma unroll
            for (int k = 0; k < TC; ++k) {
                int const e = e0 + k * NT;
                int const l = e / ncoef, r = e - l * ncoef;
                tv[k] = e < nel ? coef[lay(cl, r, line0 + l)] : 0.0;
            }
#pragma unroll
            for (int k = 0; k < TC; ++k) {
                int const e = e0 + k * NT;
                int const l = e / ncoef, r = e - l * ncoef;
                if (e < nel) tile[r * pitch + l] = tv[k];
            }
        }
    } else {
        for (int c0 = prow; c0 < ncoef; c0 += TC * ROWS) {
            double tv[TC];
#pragma unroll
            for (int k = 0; k < TC; ++k) {
                int const r = c0 + k * ROWS;
                tv[k] = r < ncoef ? coef[cb + int64_t(r) * cl.rs] : 0.0;
            }
#pragma unroll
            for (int k = 0; k < TC; ++k) {
                int const r = c0 + k * ROWS;
                if (r < ncoef) tile[r * pitch + ln] = tv[k];
            }
        }
    }
    for (int k = tid; k < ncoef * pitch; k += NT) gt[k] = 0.0;
    if (!b.uniform) {
        for (int k = tid; k < nknots; k += NT) kn[k] = b.knots[k];
        for (int k = tid; k < b.nbucket; k += NT) cells[k] = b.cells[k];
    }
    __syncthreads();
    const double* knots = b.uniform ? nullptr : kn;
    BSplinesView bl = b;
    bl.cells = b.nbucket > 0 ? cells : nullptr;
    if (active) {
        const double* tl = tile + ln;
        double* gl = gt + ln;
        for (int64_t i = prow; i < npts; i += U * ROWS) {
            double xa[U], wa[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                int64_t const ii = i + int64_t(u) * ROWS;
                xa[u] = ii < npts ? coords[pb + ii * pl.rs] : 0.0;
                wa[u] = ii < npts ? w[pb + ii * pl.rs] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                int64_t const ii = i + int64_t(u) * ROWS;
                if (ii >= npts) continue;
                double x = xa[u];
                bool outside;
                if (!apply_bc(bl, rules, x, outside)) {
                    gx[pb + ii * pl.rs] = 0.0;
                    continue;
                }
                double der[(D + 1) * 2];
                int const jmin = eval_basis_and_n_derivs<D, 1, true>(bl, knots, x, der);
                double dy = 0.0;
#pragma unroll
                for (int k = 0; k < D + 1; ++k) {
                    dy += tl[(jmin + k) * pitch] * der[k * 2 + 1];
                    atomicAdd(&gl[(jmin + k) * pitch], wa[u] * der[k * 2]);
                }
                gx[pb + ii * pl.rs] = outside ? 0.0 : wa[u] * dy;
            }
        }
    }
    __syncthreads();
    if constexpr (LINE_MAJOR) {
        for (int e = tid; e < nel; e += NT) {
            int const l = e / ncoef, r = e - l * ncoef;
            gcoef[lay(cl, r, line0 + l)] = gt[r * pitch + l];
        }
    } else if (active) {
        for (int r = prow; r < ncoef; r += ROWS) gcoef[cb + int64_t(r) * cl.rs] = gt[r * pitch + ln];
    }
}

constexpr size_t kMaxLds = 160 * 1024;

// Adjoint of the value form for batch-contiguous arrays on 16-line tiles (as
// the evaluators; eval_vjp_tiled_kernel holds two tiles and so takes 8
// lines), with one LDS tile used twice: first the coefficients, for
// gx = w * f'(x) over all the tile's points; then the coefficient gradient,
// accumulated over the same points read again and written once.  Same
// arithmetic per point as eval_vjp_tiled_kernel.
template <int D, int LB, int NT>
__global__ void __launch_bounds__(NT) eval_vjp_2pass_kernel(BSplinesView b, Rules1D rules,
                                                            const double* __restrict__ coords,
                                                            const double* __restrict__ w, double* __restrict__ gx,
                                                            Layout pl, int64_t npts, const double* __restrict__ coef,
                                                            double* __restrict__ gcoef, Layout cl, int64_t nbatch,
                                                            int ncoef, int pitch, int nknots)
{
    extern __shared__ double smem[];
    double* tile = smem;
    double* kn = tile + size_t(ncoef) * pitch;
    uint32_t* cells = reinterpret_cast<uint32_t*>(kn + nknots);
    constexpr int ROWS = NT / LB;
    constexpr int U = 4;
    constexpr int TC = 16;
    int const tid = threadIdx.x;
    int const ln = tid % LB;
    int const prow = tid / LB;
    int64_t const line0 = int64_t(blockIdx.x) * LB;
    int64_t const line = line0 + ln;
    bool const active = line < nbatch;
    int64_t const lc = active ? line : 0;
    int64_t const cb = lay(cl, 0, lc);
    int64_t const pb = lay(pl, 0, lc);
    for (int c0 = prow; c0 < ncoef; c0 += TC * ROWS) {
        double tv[TC];
#pragma unroll
        for (int k = 0; k < TC; ++k) {
            int const r = c0 + k * ROWS;
            tv[k] = r < ncoef ? coef[cb + int64_t(r) * cl.rs] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < TC; ++k) {
            int const r = c0 + k * ROWS;
            if (r < ncoef) tile[r * pitch + ln] = tv[k];
        }
    }
    if (!b.uniform) {
        for (int k = tid; k < nknots; k += NT) kn[k] = b.knots[k];
        for (int k = tid; k < b.nbucket; k += NT) cells[k] = b.cells[k];
    }
    __syncthreads();
    const double* knots = b.uniform ? nullptr : kn;
    BSplinesView bl = b;
    bl.cells = b.nbucket > 0 ? cells : nullptr;
    double* tl = tile + ln;
    // pass 1: gx from the coefficient tile
    if (active) {
        for (int64_t i = prow; i < npts; i += U * ROWS) {
            double xa[U], wa[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                int64_t const ii = i + int64_t(u) * ROWS;
                xa[u] = ii < npts ? coords[pb + ii * pl.rs] : 0.0;
                wa[u] = ii < npts ? w[pb + ii * pl.rs] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                int64_t const ii = i + int64_t(u) * ROWS;
                if (ii >= npts) continue;
                double x = xa[u];
                bool outside;
                if (!apply_bc(bl, rules, x, outside)) {
                    gx[pb + ii * pl.rs] = 0.0;
                    continue;
                }
                double der[(D + 1) * 2];
                int const jmin = eval_basis_and_n_derivs<D, 1, true>(bl, knots, x, der);
                double dy = 0.0;
#pragma unroll
                for (int k = 0; k < D + 1; ++k) dy += tl[(jmin + k) * pitch] * der[k * 2 + 1];
                gx[pb + ii * pl.rs] = outside ? 0.0 : wa[u] * dy;
            }
        }
    }
    __syncthreads();
    for (int k = tid; k < ncoef * pitch; k += NT) tile[k] = 0.0;
    __syncthreads();
    // pass 2: the coefficient gradient, the points read again
    if (active) {
        for (int64_t i = prow; i < npts; i += U * ROWS) {
            double xa[U], wa[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                int64_t const ii = i + int64_t(u) * ROWS;
                xa[u] = ii < npts ? coords[pb + ii * pl.rs] : 0.0;
                wa[u] = ii < npts ? w[pb + ii * pl.rs] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                int64_t const ii = i + int64_t(u) * ROWS;
                if (ii >= npts) continue;
                double x = xa[u];
                bool outside;
                if (!apply_bc(bl, rules, x, outside)) continue;
                double v[D + 1];
                int const jmin = eval_basis<D, true>(bl, knots, x, v);
#pragma unroll
                for (int k = 0; k < D + 1; ++k) atomicAdd(&tl[(jmin + k) * pitch], wa[u] * v[k]);
            }
        }
    }
    __syncthreads();
    if (active)
        for (int r = prow; r < ncoef; r += ROWS) gcoef[cb + int64_t(r) * cl.rs] = tile[r * pitch + ln];
}



// eval_stream_kernel for the value form (and the first derivative of uniform
// bases up to degree 3) when its shape fits: plain
// batch-contiguous points and coefficients (element (row, line) at row * rs +
// line), npts <= NST * U * ROWS, size() <= PR * ROWS, buffers below 4 GiB
// including the offsets of the point slots past the end of a line.  Uniform
// bases: 512 threads, 8 points per thread and stage, 33 prefetch rows; non-
// uniform bases (whose cell search wants more waves to hide its dependent
// LDS reads): 768 threads, 4 points per thread and stage, 22 prefetch rows,
// degrees up to 3.
template <int D, int ORDER>
bool launch_stream(BSplinesPlan const& bs, Rules1D rules, double* out, const double* coords, Layout pl,
                   int64_t npts, const double* coef, Layout cl, int64_t nbatch, hipStream_t stream)
{
    int const ncoef = bs.size();
    int const nknots = bs.view.uniform ? 0 : int(bs.knots_host.size());
    BSplinesView const b = bs.device_view();
    constexpr int LB = 16, pitch = 17;
    auto plain = [&](Layout const& L) { return L.is == 1 && L.inner >= nbatch && L.rs >= nbatch; };
    if (!plain(pl) || !plain(cl)) return false;
    // batches of fewer than two tiles per CU keep eval_tiled_kernel: nothing
    // is left to prefetch
    int64_t const ntiles = (nbatch + LB - 1) / LB;
    if (ntiles < 2 * int64_t(cu_count())) return false;
    size_t const lds = size_t(ncoef) * pitch * sizeof(double) + nknots * sizeof(double) +
                       size_t(b.nbucket) * sizeof(uint32_t);
    if (lds > kMaxLds) return false;
    unsigned const grid = unsigned(std::min<int64_t>(ntiles, cu_count()));
    // the instantiations below hold their prefetch rows and point stages in
    // registers without spilling
    auto go = [&]<int KIND, int NT, int U, int NST, int PR>() {
        constexpr int ROWS = NT / LB;
        if (npts > int64_t(NST) * U * ROWS || ncoef > PR * ROWS) return false;
        uint64_t const lim = uint64_t(1) << 32;
        // byte offsets up to the last slot of the last stage must fit 32 bits
        if ((uint64_t(NST) * U * ROWS * uint64_t(pl.rs) + uint64_t(nbatch)) * 8 >= lim) return false;
        if ((uint64_t(ncoef) * uint64_t(cl.rs) + uint64_t(nbatch)) * 8 >= lim) return false;
        uint32_t const pbytes = uint32_t((uint64_t(npts - 1) * uint64_t(pl.rs) + uint64_t(nbatch)) * 8);
        uint32_t const cbytes = uint32_t((uint64_t(ncoef - 1) * uint64_t(cl.rs) + uint64_t(nbatch)) * 8);
        auto kern = eval_stream_kernel<D, ORDER, KIND, NT, U, NST, PR>;
        hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)),
                  "hipFuncSetAttribute");
        hipLaunchKernelGGL(kern, dim3(grid), dim3(NT), lds, stream, b, rules, out, coords, uint32_t(pl.rs),
                           pbytes, coef, uint32_t(cl.rs), cbytes, nbatch, ncoef, pitch, nknots, ntiles, int(npts));
        hip_check(hipGetLastError(), "eval_stream_kernel launch");
        return true;
    };
    if constexpr (ORDER == 0) {
        if (b.uniform)
            return npts <= 2 * 8 * 32 ? go.template operator()<kUniformBasis, 512, 8, 2, 33>()
                                      : go.template operator()<kUniformBasis, 512, 8, 4, 33>();
        if constexpr (D <= 3) // higher degrees spill in this form and keep eval_tiled_kernel
            return npts <= 4 * 4 * 48 ? go.template operator()<kNonUniformBasis, 768, 4, 4, 22>()
                                      : go.template operator()<kNonUniformBasis, 768, 4, 6, 22>();
    } else if constexpr (D <= 3) {
        // the first derivative's basis takes more registers: 4 points per
        // thread and stage; uniform bases only (the non-uniform form spills)
        if (b.uniform)
            return npts <= 4 * 4 * 32 ? go.template operator()<kUniformBasis, 512, 4, 4, 33>()
                                      : go.template operator()<kUniformBasis, 512, 4, 8, 33>();
    }
    return false;
}

template <int D, int ORDER>
bool launch_orders(BSplinesPlan const& bs, Rules1D rules, double* out, const double* coords, Layout xl, Layout pl,
                   int64_t npts, const double* coef, Layout cl, int64_t nbatch, hipStream_t stream)
{
    int const ncoef = bs.size();
    int const nknots = bs.view.uniform ? 0 : int(bs.knots_host.size());
    BSplinesView const b = bs.device_view();
    size_t const extra = nknots * sizeof(double) + size_t(b.nbucket) * sizeof(uint32_t);
    auto fits = [&](int lb, int pitch) { return size_t(ncoef) * pitch * sizeof(double) + extra <= kMaxLds; };
    bool const line_major = pl.rs == 1;
    auto go = [&]<int LB, int NT>(int pitch) {
        size_t const lds = size_t(ncoef) * pitch * sizeof(double) + extra;
        auto kern = b.uniform ? (line_major ? eval_tiled_kernel<D, ORDER, LB, NT, true, kUniformBasis>
                                            : eval_tiled_kernel<D, ORDER, LB, NT, false, kUniformBasis>)
                              : (line_major ? eval_tiled_kernel<D, ORDER, LB, NT, true, kNonUniformBasis>
                                            : eval_tiled_kernel<D, ORDER, LB, NT, false, kNonUniformBasis>);
        hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)),
                  "hipFuncSetAttribute");
        int64_t const blocks = (nbatch + LB - 1) / LB;
        hipLaunchKernelGGL(kern, dim3(unsigned(blocks)), dim3(NT), lds, stream, b, rules, out, coords, xl, pl, npts,
                           coef, cl, nbatch, ncoef, pitch, nknots);
        hip_check(hipGetLastError(), "eval_tiled_kernel launch");
        return true;
    };
    if (ncoef > 2560) return false;
    auto same = [](Layout const& a, Layout const& c) {
        return a.rs == c.rs && a.inner == c.inner && a.is == c.is && a.os == c.os;
    };
    if constexpr (ORDER <= 1) {
        if (same(xl, pl) && launch_stream<D, ORDER>(bs, rules, out, coords, pl, npts, coef, cl, nbatch, stream))
            return true;
    }
    if (fits(16, 17)) return go.template operator()<16, 1024>(17);
    if (fits(16, 16)) return go.template operator()<16, 1024>(16);
    if (fits(8, 9)) return go.template operator()<8, 1024>(9);
    if (fits(8, 8)) return go.template operator()<8, 1024>(8);
    return false;
}

} // namespace

int cu_count()
{
    static int cached[64] = {};
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= 64) dev = 0;
    if (cached[dev] == 0) {
        int n = 0;
        hip_check(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
        cached[dev] = n > 0 ? n : 1;
    }
    return cached[dev];
}

bool launch_eval_tiled(BSplinesPlan const& bs, Rules1D rules, int order, double* out, const double* coords,
                       Layout xl, Layout pl, int64_t npts, const double* coef, Layout cl, int64_t nbatch,
                       hipStream_t stream)
{
    // tiny batches, and mixed layouts (points and coefficients not both
    // batch-contiguous or both interpolation-contiguous), take the generic kernel
    if (npts < 64 || nbatch < 8) return false;
    if ((pl.rs == 1) != (cl.rs == 1)) return false;
    // 32-bit point indices and point strides in the kernel
    if (npts >= (int64_t(1) << 30) || pl.rs >= (int64_t(1) << 32) || xl.rs >= (int64_t(1) << 32)) return false;
    bool launched = false;
    dispatch_degree(bs.view.degree, [&]<int D>() {
        auto ord = [&]<int O>() {
            if constexpr (O <= D) {
                if (!launched && order == O)
                    launched = launch_orders<D, O>(bs, rules, out, coords, xl, pl, npts, coef, cl, nbatch, stream);
            }
        };
        ord.template operator()<0>();
        ord.template operator()<1>();
        ord.template operator()<2>();
        ord.template operator()<3>();
    });
    return launched;
}

bool launch_eval_vjp_tiled(BSplinesPlan const& bs, Rules1D rules, const double* coords, const double* w,
                           double* gx, Layout pl, int64_t npts, const double* coef, double* gcoef, Layout cl,
                           int64_t nbatch, hipStream_t stream)
{
    if (npts < 64 || nbatch < 8) return false;
    if ((pl.rs == 1) != (cl.rs == 1)) return false;
    int const ncoef = bs.size();
    if (ncoef > 2560) return false;
    int const nknots = bs.view.uniform ? 0 : int(bs.knots_host.size());
    BSplinesView const b = bs.device_view();
    size_t const extra = nknots * sizeof(double) + size_t(b.nbucket) * sizeof(uint32_t);
    auto lds_of = [&](int pitch) { return 2 * size_t(ncoef) * pitch * sizeof(double) + extra; };
    int pitch = 0;
    if (lds_of(9) <= kMaxLds) pitch = 9;
    else if (lds_of(8) <= kMaxLds) pitch = 8;
    else return false;
    bool const line_major = pl.rs == 1;
    bool launched = false;
    if (!line_major && nbatch >= 16 * 2 * int64_t(cu_count()) && b.uniform && b.degree <= 3) {
        // batch-contiguous, uniform bases up to degree 3, at least two tiles
        // per CU: 16-line tiles, one LDS tile used for both passes (measured
        // faster than the 8-line kernel there; with the dearer basis of
        // non-uniform or quintic splines, computing it twice costs more than
        // the wider tile gains, and small batches want the 8-line kernel's
        // twice as many workgroups)
        size_t const lds2 = size_t(ncoef) * 17 * sizeof(double) + extra;
        if (lds2 <= kMaxLds) {
            dispatch_degree(b.degree, [&]<int D>() {
                constexpr int LB = 16, NT = 1024;
                auto kern = eval_vjp_2pass_kernel<D, LB, NT>;
                hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                              hipFuncAttributeMaxDynamicSharedMemorySize, int(lds2)),
                          "hipFuncSetAttribute");
                int64_t const blocks = (nbatch + LB - 1) / LB;
                hipLaunchKernelGGL(kern, dim3(unsigned(blocks)), dim3(NT), lds2, stream, b, rules, coords, w, gx, pl,
                                   npts, coef, gcoef, cl, nbatch, ncoef, 17, nknots);
                hip_check(hipGetLastError(), "eval_vjp_2pass_kernel launch");
                launched = true;
            });
            if (launched) return true;
        }
    }
    dispatch_degree(b.degree, [&]<int D>() {
        constexpr int LB = 8, NT = 1024;
        auto kern = line_major ? eval_vjp_tiled_kernel<D, LB, NT, true> : eval_vjp_tiled_kernel<D, LB, NT, false>;
        size_t const lds = lds_of(pitch);
        hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)),
                  "hipFuncSetAttribute");
        int64_t const blocks = (nbatch + LB - 1) / LB;
        hipLaunchKernelGGL(kern, dim3(unsigned(blocks)), dim3(NT), lds, stream, b, rules, coords, w, gx, pl,
                           npts, coef, gcoef, cl, nbatch, ncoef, pitch, nknots);
        hip_check(hipGetLastError(), "eval_vjp_tiled_kernel launch");
        launched = true;
    });
    return launched;
}

} // namespace ddcx
