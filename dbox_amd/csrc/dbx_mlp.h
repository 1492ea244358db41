// dbx_mlp.h -- learned warm start for new contact points (host side).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/dbox_hip.h"
#include "dbx_device.h"
#include "dbx_mlp_dev.h"

namespace dbx {

struct MlpState {
    bool ready = false;     // parameters and normalisation loaded
    float *params = nullptr; // fp32 parameters (kMlpParams)
    float *grad = nullptr;
    float *adam1 = nullptr, *adam2 = nullptr;
    float *norm = nullptr;   // kMlpNorm
    bf16x8 *frag = nullptr;  // packed MFMA operands
    float *loss = nullptr;
    int adamStep = 0;
    std::vector<void *> allocs;
};

void mlp_release(MlpState &m);
hipError_t mlp_ensure(MlpState &m, hipStream_t s);
MlpDev mlp_dev(const MlpState &m);
// Uploads parameters and/or normalisation (either may be null) and repacks.
hipError_t mlp_upload(MlpState &m, const float *params, const float *norm, hipStream_t s);
// y[2i], y[2i+1] = predicted normal and tangent impulse of feature row i (device pointers).
hipError_t mlp_forward(MlpState &m, const float *dX, int n, float *dY, hipStream_t s);
// Accumulates into m.grad the gradient of the mean squared error over `batch`
// samples (rows drawn with replacement by `seed`, or rows 0..batch-1 if seed is 0)
// and adds the loss to *dLoss (a zeroed device float), or writes it to m.loss if dLoss is null.
hipError_t mlp_gradients(MlpState &m, const float *dX, const float *dT, long long n, long long batch, uint64_t seed,
                         float *dLoss, hipStream_t s);
hipError_t mlp_fill(float *p, long long n, float v, hipStream_t s);
hipError_t mlp_adam_step(MlpState &m, float lr, hipStream_t s);
hipError_t mlp_clear_grad(MlpState &m, hipStream_t s);

} // namespace dbx
