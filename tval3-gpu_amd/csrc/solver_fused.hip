// solver_fused.hip -- placeholder until the fused solver lands.
#include "tv3_internal.hpp"

namespace tv3 {
void solve_fused(tv3_ctx *, tv3_op *, const tv3_opts &, const int32_t *, const float *, float *,
                 const float *, tv3_info *, tv3_history *) {
    throw tv3_error(TV3_E_UNSUPPORTED, "fused solver not built yet");
}
}  // namespace tv3
