// mlp.hip -- FP32 cost model for per-query dispatch (forward, backward, Adam).
#include <hip/hip_runtime.h>

#include "engine.h"

namespace ab {

ab_status mlp_decide(ab_ctx *c, const float *d_feat, uint32_t nq, uint8_t *d_sel)
{
    (void)c;
    (void)d_feat;
    (void)nq;
    (void)d_sel;
    set_error("cost model not available");
    return AB_UNSUPPORTED;
}

int mlp_decide_host(const ab_ctx *c, const float *f)
{
    (void)c;
    (void)f;
    return AB_PATH_SPARSE;
}

}  // namespace ab
