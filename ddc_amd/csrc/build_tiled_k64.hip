// FP64 tiled build on the 64 x 32 tiling (main blocks of up to 2048 rows).
#include "build_tiled_kernel.hpp"

namespace ddcx {

bool launch_build_tiled_k64(BuilderPlan const& p, double* coef, Layout cl, const double* vals, Layout vl,
                            const double* dl, const double* dr, Layout dlay, int64_t nbatch, hipStream_t stream)
{
    return tiled_build::launch_K<64, 32>(p, coef, cl, vals, vl, dl, dr, dlay, nbatch, stream);
}

} // namespace ddcx
