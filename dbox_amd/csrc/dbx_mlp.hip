// dbx_mlp.hip -- learned warm start: parameter packing, batched forward pass,
// gradients and Adam training on the GPU, and the training-sample log.
//
// Nothing here has a counterpart in the reference.  The network predicts the
// normal and tangent impulse that the plain solver converges to for a contact
// point that b2Contact.Update (b2contact.d:306-324) could not match to a point
// of the previous manifold and therefore starts at zero.
#include "dbx_mlp.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace dbx {

// ---------------------------------------------------------------------------
// Fragment packing (fp32 parameters -> bf16 MFMA A operands).
// ---------------------------------------------------------------------------
__global__ void k_mlp_pack(const float *__restrict__ P, bf16x8 *__restrict__ frag) {
    const int f = blockIdx.x, lane = threadIdx.x, i = lane & 15, g = lane >> 4;
    float v[8];
    for (int j = 0; j < 8; ++j) {
        float x = 0.0f;
        if (f < kFragW2) { // W1 tile t: rows 16t + i, inputs 8g + j
            const int t = f - kFragW1, in = 8 * g + j;
            if (in < kMlpIn) x = P[kOffW1 + (16 * t + i) * kMlpIn + in];
        } else if (f < kFragW3) { // W2 tile t, K step s
            const int t = (f - kFragW2) >> 1, s = (f - kFragW2) & 1;
            x = P[kOffW2 + (16 * t + i) * kMlpHidden + mlp_kappa(s, g, j)];
        } else if (f < kFragW2T) { // W3, K step s; rows are the two outputs
            const int s = f - kFragW3;
            if (i < kMlpOut) x = P[kOffW3 + i * kMlpHidden + mlp_kappa(s, g, j)];
        } else if (f < kFragW3T) { // W2^T tile tk (rows = W2 inputs), K step s over W2 outputs
            const int tk = (f - kFragW2T) >> 1, s = (f - kFragW2T) & 1;
            x = P[kOffW2 + mlp_kappa(s, g, j) * kMlpHidden + 16 * tk + i];
        } else { // W3^T tile tk, one K step whose slots j < 4 of group g are output 4g + j
            const int tk = f - kFragW3T, o = 4 * g + j;
            if (j < 4 && o < kMlpOut) x = P[kOffW3 + o * kMlpHidden + 16 * tk + i];
        }
        v[j] = x;
    }
    frag[64 * f + lane] = bf8_from(v);
}

// ---------------------------------------------------------------------------
// Batched forward pass: x rows of kMlpFeatStride raw features -> impulses.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float2 mlp_output(const MlpDev &m, const f32x4 &y) {
    return make_float2(y[0] * m.norm[kNormOutScale] + m.norm[kNormOutMean],
                       y[1] * m.norm[kNormOutScale + 1] + m.norm[kNormOutMean + 1]);
}

__global__ __launch_bounds__(64) void k_mlp_forward(MlpDev m, const float *__restrict__ X, int n, float *__restrict__ Y) {
    const int lane = threadIdx.x, g = lane >> 4, col = lane & 15;
    for (int base = 16 * blockIdx.x; base < n; base += 16 * gridDim.x) {
        const int p = base + col;
        const bf16x8 x = mlp_input_operand(m, p < n ? X + (size_t)p * kMlpFeatStride : nullptr, g);
        MlpFwd F;
        mlp_forward16(m, x, F);
        if (g == 0 && p < n) {
            float2 o = mlp_output(m, F.y);
            Y[2 * (size_t)p] = o.x;
            Y[2 * (size_t)p + 1] = o.y;
        }
    }
}

// ---------------------------------------------------------------------------
// Gradients of the mean squared error of the normalised outputs,
//   L = 1/B sum_p sum_o (y_o(p) - t_o(p))^2,  t_o = (label_o - mean_o) / scale_o,
// over B samples; sample b of the batch is row sample_index(b) of X / T.
// One wave per workgroup handles 32 samples per round: two forward/backward
// passes of 16 points in registers, then the weight gradients as MFMA products
// over the 32 samples, with the activations staged feature-major in LDS.
// ---------------------------------------------------------------------------
constexpr int kRowPad = 40; // bf16 per staged row (32 samples + padding against bank conflicts)

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
__device__ __forceinline__ long long sample_index(uint64_t seed, long long b, long long n) {
    return seed == 0 ? b : (long long)(mix64(seed ^ (uint64_t)b * 0xD1B54A32D192ED03ull) % (uint64_t)n);
}

__device__ __forceinline__ bf16x8 lds_row8(const __bf16 *row, int g) {
    bf16x8 r;
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] = row[8 * g + j];
    return r;
}

__global__ __launch_bounds__(64) void k_mlp_grad(MlpDev m, const float *__restrict__ X, const float *__restrict__ T,
                                                 long long n, long long B, uint64_t seed, float *__restrict__ grad,
                                                 float *__restrict__ loss) {
    __shared__ __bf16 xs[16 * kRowPad], dys[16 * kRowPad];
    __shared__ __bf16 h1s[64 * kRowPad], dz1s[64 * kRowPad], h2s[64 * kRowPad], dz2s[64 * kRowPad];
    const int lane = threadIdx.x, g = lane >> 4, col = lane & 15;
    const bf16x8 *fr = m.frag + lane;
    const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
    f32x4 a1[4], a2[4][4], a3[4];
    for (int t = 0; t < 4; ++t) {
        a1[t] = zero;
        a3[t] = zero;
        for (int u = 0; u < 4; ++u) a2[t][u] = zero;
    }
    float gb1 = 0.0f, gb2 = 0.0f, gb3 = 0.0f, lossAcc = 0.0f;
    const float invB = 1.0f / (float)B;
    for (int r = lane; r < 16 * kRowPad; r += 64) dys[r] = (__bf16)0.0f;
    const long long rounds = (B + 31) / 32;
    for (long long rd = blockIdx.x; rd < rounds; rd += gridDim.x) {
        __syncthreads();
        for (int q = 0; q < 2; ++q) {
            const long long b = 32 * rd + 16 * q + col;
            const bool valid = b < B;
            const long long idx = valid ? sample_index(seed, b, n) : 0;
            const float *xr = valid ? X + idx * kMlpFeatStride : nullptr;
            const bf16x8 x = mlp_input_operand(m, xr, g);
            const int pc = 16 * q + col; // column of this point in the staged rows
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (8 * g + j < 16) xs[(8 * g + j) * kRowPad + pc] = x[j];
            MlpFwd F;
            mlp_forward16(m, x, F);
            // output error (lanes of group 0 hold outputs 0 and 1)
            float dy0 = 0.0f, dy1 = 0.0f;
            if (g == 0 && valid) {
                const float t0 = (T[2 * idx] - m.norm[kNormOutMean]) / m.norm[kNormOutScale];
                const float t1 = (T[2 * idx + 1] - m.norm[kNormOutMean + 1]) / m.norm[kNormOutScale + 1];
                const float e0 = F.y[0] - t0, e1 = F.y[1] - t1;
                lossAcc += e0 * e0 + e1 * e1;
                dy0 = 2.0f * e0 * invB;
                dy1 = 2.0f * e1 * invB;
            }
            if (g == 0) {
                dys[0 * kRowPad + pc] = (__bf16)dy0;
                dys[1 * kRowPad + pc] = (__bf16)dy1;
            }
            bf16x8 dyop;
#pragma unroll
            for (int j = 0; j < 8; ++j) dyop[j] = (__bf16)0.0f;
            dyop[0] = (__bf16)dy0;
            dyop[1] = (__bf16)dy1;
            // dZ2 = (W3^T dY) * [z2 > 0]
            f32x4 dz2[4], h[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                f32x4 dh = mfma(fr[64 * (kFragW3T + t)], dyop, zero);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int f = 16 * t + 4 * g + i;
                    dz2[t][i] = F.z2[t][i] > 0.0f ? dh[i] : 0.0f;
                    h[t][i] = F.z2[t][i] > 0.0f ? F.z2[t][i] : 0.0f;
                    h2s[f * kRowPad + pc] = (__bf16)h[t][i];
                    dz2s[f * kRowPad + pc] = (__bf16)dz2[t][i];
                }
            }
            // dZ1 = (W2^T dZ2) * [z1 > 0]
            const bf16x8 d0 = mlp_b_from_tiles(dz2, 0), d1 = mlp_b_from_tiles(dz2, 1);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                f32x4 dh = mfma(fr[64 * (kFragW2T + 2 * t)], d0, zero);
                dh = mfma(fr[64 * (kFragW2T + 2 * t + 1)], d1, dh);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int f = 16 * t + 4 * g + i;
                    const float z = F.z1[t][i];
                    h1s[f * kRowPad + pc] = (__bf16)(z > 0.0f ? z : 0.0f);
                    dz1s[f * kRowPad + pc] = (__bf16)(z > 0.0f ? dh[i] : 0.0f);
                }
            }
        }
        __syncthreads();
        // weight gradients over the 32 samples (summation label: sample 8g + j)
#pragma unroll
        for (int to = 0; to < 4; ++to) {
            const bf16x8 A1 = lds_row8(dz1s + (16 * to + col) * kRowPad, g);
            a1[to] = mfma(A1, lds_row8(xs + col * kRowPad, g), a1[to]);
            const bf16x8 A2 = lds_row8(dz2s + (16 * to + col) * kRowPad, g);
#pragma unroll
            for (int tk = 0; tk < 4; ++tk) a2[to][tk] = mfma(A2, lds_row8(h1s + (16 * tk + col) * kRowPad, g), a2[to][tk]);
            a3[to] = mfma(lds_row8(dys + col * kRowPad, g), lds_row8(h2s + (16 * to + col) * kRowPad, g), a3[to]);
        }
        // bias gradients: lane f sums row f over the 32 samples
        for (int p = 0; p < 32; ++p) {
            gb1 += (float)dz1s[lane * kRowPad + p];
            gb2 += (float)dz2s[lane * kRowPad + p];
            if (lane < kMlpOut) gb3 += (float)dys[lane * kRowPad + p];
        }
    }
    // D row 4g + i, column col of each tile
#pragma unroll
    for (int to = 0; to < 4; ++to) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int o = 16 * to + 4 * g + i;
            if (col < kMlpIn) atomicAdd(&grad[kOffW1 + o * kMlpIn + col], a1[to][i]);
#pragma unroll
            for (int tk = 0; tk < 4; ++tk) atomicAdd(&grad[kOffW2 + o * kMlpHidden + 16 * tk + col], a2[to][tk][i]);
            const int o3 = 4 * g + i;
            if (o3 < kMlpOut) atomicAdd(&grad[kOffW3 + o3 * kMlpHidden + 16 * to + col], a3[to][i]);
        }
    }
    atomicAdd(&grad[kOffB1 + lane], gb1);
    atomicAdd(&grad[kOffB2 + lane], gb2);
    if (lane < kMlpOut) atomicAdd(&grad[kOffB3 + lane], gb3);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) lossAcc += __shfl_xor(lossAcc, o, 64);
    if (lane == 0 && loss) atomicAdd(loss, lossAcc * invB);
}

// Adam (Kingma & Ba 2015) on the fp32 parameters; clears the gradient.
__global__ void k_mlp_adam(float *__restrict__ P, float *__restrict__ G, float *__restrict__ M1, float *__restrict__ M2,
                           float lr, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kMlpParams) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float gi = G[i];
    const float m = b1 * M1[i] + (1.0f - b1) * gi;
    const float v = b2 * M2[i] + (1.0f - b2) * gi * gi;
    M1[i] = m;
    M2[i] = v;
    P[i] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
    G[i] = 0.0f;
}

__global__ void k_fill(float *p, long long n, float v) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
void mlp_release(MlpState &m) {
    for (void *p : m.allocs) (void)hipFree(p);
    m.allocs.clear();
    m = MlpState{};
}

static hipError_t mlp_alloc(MlpState &m, void **p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes < 16 ? 16 : bytes);
    if (e == hipSuccess) m.allocs.push_back(*p);
    return e;
}

hipError_t mlp_ensure(MlpState &m, hipStream_t s) {
    if (m.params) return hipSuccess;
    hipError_t e;
    if ((e = mlp_alloc(m, (void **)&m.params, kMlpParams * sizeof(float))) != hipSuccess) return e;
    if ((e = mlp_alloc(m, (void **)&m.grad, kMlpParams * sizeof(float))) != hipSuccess) return e;
    if ((e = mlp_alloc(m, (void **)&m.adam1, kMlpParams * sizeof(float))) != hipSuccess) return e;
    if ((e = mlp_alloc(m, (void **)&m.adam2, kMlpParams * sizeof(float))) != hipSuccess) return e;
    if ((e = mlp_alloc(m, (void **)&m.norm, kMlpNorm * sizeof(float))) != hipSuccess) return e;
    if ((e = mlp_alloc(m, (void **)&m.frag, kMlpFrags * 64 * sizeof(bf16x8))) != hipSuccess) return e;
    if ((e = mlp_alloc(m, (void **)&m.loss, sizeof(float))) != hipSuccess) return e;
    hipLaunchKernelGGL(k_fill, dim3(8), dim3(256), 0, s, m.grad, (long long)kMlpParams, 0.0f);
    hipLaunchKernelGGL(k_fill, dim3(8), dim3(256), 0, s, m.adam1, (long long)kMlpParams, 0.0f);
    hipLaunchKernelGGL(k_fill, dim3(8), dim3(256), 0, s, m.adam2, (long long)kMlpParams, 0.0f);
    return hipGetLastError();
}

MlpDev mlp_dev(const MlpState &m) {
    MlpDev d;
    d.frag = m.frag;
    d.params = m.params;
    d.norm = m.norm;
    return d;
}

hipError_t mlp_upload(MlpState &m, const float *params, const float *norm, hipStream_t s) {
    hipError_t e = mlp_ensure(m, s);
    if (e != hipSuccess) return e;
    if (params && (e = hipMemcpyAsync(m.params, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice, s)) != hipSuccess)
        return e;
    if (norm && (e = hipMemcpyAsync(m.norm, norm, kMlpNorm * sizeof(float), hipMemcpyHostToDevice, s)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_mlp_pack, dim3(kMlpFrags), dim3(64), 0, s, m.params, m.frag);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // a new network restarts the optimiser
    hipLaunchKernelGGL(k_fill, dim3(8), dim3(256), 0, s, m.adam1, (long long)kMlpParams, 0.0f);
    hipLaunchKernelGGL(k_fill, dim3(8), dim3(256), 0, s, m.adam2, (long long)kMlpParams, 0.0f);
    m.adamStep = 0;
    m.ready = true;
    return hipGetLastError();
}

hipError_t mlp_forward(MlpState &m, const float *dX, int n, float *dY, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int groups = (n + 15) / 16;
    hipLaunchKernelGGL(k_mlp_forward, dim3(groups < 4096 ? groups : 4096), dim3(64), 0, s, mlp_dev(m), dX, n, dY);
    return hipGetLastError();
}

hipError_t mlp_gradients(MlpState &m, const float *dX, const float *dT, long long n, long long batch, uint64_t seed,
                         float *dLoss, hipStream_t s) {
    if (!dLoss) {
        hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, s, m.loss, 1LL, 0.0f);
        dLoss = m.loss;
    }
    const long long rounds = (batch + 31) / 32;
    const long long waves = rounds < 2048 ? rounds : 2048;
    hipLaunchKernelGGL(k_mlp_grad, dim3((unsigned)waves), dim3(64), 0, s, mlp_dev(m), dX, dT, n, batch, seed, m.grad,
                       dLoss);
    return hipGetLastError();
}

hipError_t mlp_fill(float *p, long long n, float v, hipStream_t s) {
    hipLaunchKernelGGL(k_fill, dim3(64), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t mlp_adam_step(MlpState &m, float lr, hipStream_t s) {
    m.adamStep++;
    const float c1 = 1.0f - std::pow(0.9f, (float)m.adamStep), c2 = 1.0f - std::pow(0.999f, (float)m.adamStep);
    hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, m.params, m.grad, m.adam1, m.adam2,
                       lr, c1, c2);
    hipLaunchKernelGGL(k_mlp_pack, dim3(kMlpFrags), dim3(64), 0, s, m.params, m.frag);
    return hipGetLastError();
}

hipError_t mlp_clear_grad(MlpState &m, hipStream_t s) {
    hipLaunchKernelGGL(k_fill, dim3(8), dim3(256), 0, s, m.grad, (long long)kMlpParams, 0.0f);
    return hipGetLastError();
}

} // namespace dbx
