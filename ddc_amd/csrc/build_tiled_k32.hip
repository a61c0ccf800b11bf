// FP64 tiled build on the 32 x 32 tiling (main blocks of 256 to 1024 rows)
// and the launcher that picks the tiling.
#include "build_tiled_kernel.hpp"

namespace ddcx {

bool launch_build_tiled_k64(BuilderPlan const& p, double* coef, Layout cl, const double* vals, Layout vl,
                            const double* dl, const double* dr, Layout dlay, int64_t nbatch, hipStream_t stream);
bool launch_build_tiled_short(BuilderPlan const& p, double* coef, Layout cl, const double* vals, Layout vl,
                              const double* dl, const double* dr, Layout dlay, int64_t nbatch, hipStream_t stream);

bool launch_build_tiled(BuilderPlan const& p, double* coef, Layout cl, const double* vals, Layout vl,
                        const double* dl, const double* dr, Layout dlay, int64_t nbatch, hipStream_t stream)
{
    if (!p.tiled.ok) return false;
    if (!tiled::offsets_fit(vl, p.tiled.K * p.tiled.L + p.tview.q0, nbatch) ||
        !tiled::offsets_fit(cl, p.view.size, nbatch))
        return false;
    if (p.tiled.K == 32 && p.tiled.L == 32) {
        // fewer than four 32-line tiles per CU: 512-thread workgroups of 16
        // lines, two resident per CU, so that one's loads and stores overlap
        // the other's sweeps
        if (nbatch < 4 * 32 * int64_t(cu_count()))
            return tiled_build::launch_K<32, 32, 512>(p, coef, cl, vals, vl, dl, dr, dlay, nbatch, stream);
        return tiled_build::launch_K<32, 32>(p, coef, cl, vals, vl, dl, dr, dlay, nbatch, stream);
    }
    if (p.tiled.K == 64 && p.tiled.L == 32)
        return launch_build_tiled_k64(p, coef, cl, vals, vl, dl, dr, dlay, nbatch, stream);
    if ((p.tiled.K == 32 && p.tiled.L == 8) || p.tiled.K == 16)
        return launch_build_tiled_short(p, coef, cl, vals, vl, dl, dr, dlay, nbatch, stream);
    return false;
}

} // namespace ddcx
