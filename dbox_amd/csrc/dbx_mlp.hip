// dbx_mlp.hip -- learned warm start for new contact points.
#include "dbx_mlp.h"

namespace dbx {

void mlp_release(MlpState &m) { m.ready = false; }

dbx_status mlp_step_hook(MlpState &m, const Dev &, int, const StepParams &, int, hipStream_t) {
    return m.ready ? DBX_OK : DBX_E_STATE;
}

} // namespace dbx
