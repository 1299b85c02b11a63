// Plain FP64 form with the reference's rounding: built with
// -ffp-contract=off -DHOWMHD_EXACT=1 (true divisions, reference evaluation
// order), so every value equals the one a reference CPU run built without
// FMA contraction computes.  Also the FP64 sample collection of the learned
// indicator (labels from these WENO weights).
#include "stage.hpp"

#if !HOWMHD_EXACT
#error "step_exact.hip is built with HOWMHD_EXACT=1"
#endif

namespace howmhd {

int stage_fp64(howmhd_gpu_ctx* c, int iw, double dt) { return launch_stage<double>(c, iw, dt); }

int collect_fp64(howmhd_gpu_ctx* c, int dir, float* X, float* y, float* logits) {
    SweepArgs<double> a = sweep_args<double>(c, dir, c->qbuf[c->qid[0]]);
    a.feat = X;
    a.label = y;
    a.logit = logits;
    a.stats = nullptr;
    launch_sweep<double, SW_COLLECT>(c, dir, a);
    HCHECK(c, hipGetLastError());
    return HOWMHD_OK;
}

}  // namespace howmhd
