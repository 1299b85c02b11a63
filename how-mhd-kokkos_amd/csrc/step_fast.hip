// FMA forms: built with -ffp-contract=fast -DHOWMHD_EXACT=0 (reciprocal
// forms for constant divisors and the WENO weight normalisation).  fp64fast
// runs the FP64 method with this rounding; fp32 runs the sweeps and CT EMFs
// in FP32.
#include "stage.hpp"

#if HOWMHD_EXACT
#error "step_fast.hip is built with HOWMHD_EXACT=0"
#endif

namespace howmhd {

int stage_fp64fast(howmhd_gpu_ctx* c, int iw, double dt) { return launch_stage<double>(c, iw, dt); }
int stage_fp32(howmhd_gpu_ctx* c, int iw, double dt) { return launch_stage<float>(c, iw, dt); }

}  // namespace howmhd
