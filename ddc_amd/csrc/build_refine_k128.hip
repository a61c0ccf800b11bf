// Mixed-precision build on the 128 x 16 tiling (main blocks of up to 2048 rows).
#include "build_refine_kernel.hpp"

namespace ddcx {

bool launch_refine_k128(BuilderPlan const& p, double* coef, Layout cl, const double* vals, Layout vl,
                        const double* dl, const double* dr, Layout dlay, int64_t nbatch, hipStream_t stream)
{
    return refine::launch_K<128, 16>(p, coef, cl, vals, vl, dl, dr, dlay, nbatch, stream);
}

} // namespace ddcx
