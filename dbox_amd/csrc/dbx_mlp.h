// dbx_mlp.h -- learned warm start for new contact points.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/dbox_hip.h"
#include "dbx_device.h"

namespace dbx {

struct MlpState {
    bool ready = false;
};

void mlp_release(MlpState &m);
dbx_status mlp_step_hook(MlpState &m, const Dev &d, int cur, const StepParams &sp, int block, hipStream_t s);

} // namespace dbx
