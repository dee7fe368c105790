// Learned initial correction (placeholder until the network kernels land).
#include "cd_internal.h"

int cd_learned_predict(cd_context* ctx, const float*, const float*, float*, int, int) {
    if (ctx) ctx->set_error("no learned-init network attached");
    return CD_ERR_STATE;
}
