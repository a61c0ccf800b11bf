// FP64 tiled build for short lines (main blocks of 64 to 255 rows, such as
// the passes of a 3D build on a 128^3 grid), 32 lines per workgroup as on
// the 32 x 32 tiling, so that the staging of interpolation-contiguous tiles
// fits in LDS as well: up to 144 rows on 16 segments of 8 or 9 rows (512
// threads, so that every thread holds rows of the line), longer lines on 32
// segments of 8 rows (1024 threads).
#include "build_tiled_kernel.hpp"

namespace ddcx {

bool launch_build_tiled_short(BuilderPlan const& p, double* coef, Layout cl, const double* vals, Layout vl,
                              const double* dl, const double* dr, Layout dlay, int64_t nbatch, hipStream_t stream)
{
    if (p.tiled.K == 16 && p.tiled.L == 8)
        return tiled_build::launch_K<16, 8, 512>(p, coef, cl, vals, vl, dl, dr, dlay, nbatch, stream);
    if (p.tiled.K == 16 && p.tiled.L == 9)
        return tiled_build::launch_K<16, 9, 512>(p, coef, cl, vals, vl, dl, dr, dlay, nbatch, stream);
    return tiled_build::launch_K<32, 8>(p, coef, cl, vals, vl, dl, dr, dlay, nbatch, stream);
}

} // namespace ddcx
