// FP64 tiled build on the 32 x 8 tiling: main blocks of 64 to 255 rows
// (short lines, such as the passes of a 3D build on a 128^3 grid), 32 lines
// per workgroup as on the 32 x 32 tiling, so that the staging of
// interpolation-contiguous tiles fits in LDS as well.
#include "build_tiled_kernel.hpp"

namespace ddcx {

bool launch_build_tiled_short(BuilderPlan const& p, double* coef, Layout cl, const double* vals, Layout vl,
                              const double* dl, const double* dr, Layout dlay, int64_t nbatch, hipStream_t stream)
{
    return tiled_build::launch_K<32, 8>(p, coef, cl, vals, vl, dl, dr, dlay, nbatch, stream);
}

} // namespace ddcx
