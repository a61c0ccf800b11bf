// Mixed-precision build on the 64 x 16 tiling (main blocks of up to 1024
// rows), and the launcher that picks the tiling.
#include "build_refine_kernel.hpp"

namespace ddcx {

bool launch_refine_k128(BuilderPlan const& p, double* coef, Layout cl, const double* vals, Layout vl,
                        const double* dl, const double* dr, Layout dlay, int64_t nbatch, hipStream_t stream);

bool launch_build_refine(BuilderPlan const& p, double* coef, Layout cl, const double* vals, Layout vl,
                         const double* dl, const double* dr, Layout dlay, int64_t nbatch, hipStream_t stream)
{
    if (!p.refine.ok) return false;
    TiledHost const& t = p.refine.t;
    if (!tiled::offsets_fit(vl, t.K * t.L + p.rview.q0, nbatch) || !tiled::offsets_fit(cl, p.view.size, nbatch))
        return false;
    if (t.K == 64 && t.L == 16) return refine::launch_K<64, 16>(p, coef, cl, vals, vl, dl, dr, dlay, nbatch, stream);
    if (t.K == 128 && t.L == 16) return launch_refine_k128(p, coef, cl, vals, vl, dl, dr, dlay, nbatch, stream);
    return false;
}

} // namespace ddcx
