// dbx_mlp_dev.h -- device side of the learned warm start: a 14 -> 64 -> 64 -> 2
// ReLU network whose products run on the bf16 matrix cores of CDNA4
// (v_mfma_f32_16x16x32_bf16, fp32 accumulation).
//
// Layout.  Activations are kept feature-major: one MFMA computes
// D[out feature][point] = W[out][in] * X[in][point] for 16 output features and
// 16 points, so a wave evaluates 16 points per MFMA column.  The MFMA places
// row r = 4 * (lane / 16) + i, column lane % 16 of D in lane `lane`, element i;
// operand A row lane % 16 and operand B column lane % 16 hold eight values of the
// summation index each.  Which summation index a (lane group, slot) pair stands
// for does not matter as long as A and B agree, so each layer's B operand is
// assembled from the previous layer's D tiles without moving data between
// lanes: slot j of lane group g carries feature kappa(s, g, j) of K step s,
//   kappa(s, g, j) = 32 s + (j < 4 ? 4 g + j : 16 + 4 g + j - 4),
// and the weight fragments are packed with the same labelling.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dbx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kMlpIn = 14, kMlpHidden = 64, kMlpOut = 2;
// fp32 parameter vector: W1[64][14], b1[64], W2[64][64], b2[64], W3[2][64], b3[2]
constexpr int kOffW1 = 0, kOffB1 = 896, kOffW2 = 960, kOffB2 = 5056, kOffW3 = 5120, kOffB3 = 5248;
constexpr int kMlpParams = 5250;
// normalisation: input mean[14], input scale[14], output mean[2], output scale[2]
constexpr int kMlpNorm = 32;
constexpr int kNormInMean = 0, kNormInScale = 14, kNormOutMean = 28, kNormOutScale = 30;
// packed A fragments, 64 lanes each: forward W1 (4 tiles), W2 (4 tiles x 2 K steps),
// W3 (2 K steps); backward W2^T (4 x 2), W3^T (4 tiles, one K step)
constexpr int kFragW1 = 0, kFragW2 = 4, kFragW3 = 12, kFragW2T = 14, kFragW3T = 22, kMlpFrags = 26;
constexpr int kMlpFeatStride = 16; // floats per logged / staged feature row (14 used)

struct MlpDev {
    const bf16x8 *frag = nullptr;  // [kMlpFrags][64]
    const float *params = nullptr; // fp32 parameters (biases are read from here)
    const float *norm = nullptr;   // kMlpNorm floats
};

__device__ __forceinline__ int mlp_kappa(int s, int g, int j) { return 32 * s + (j < 4 ? 4 * g + j : 12 + 4 * g + j); }

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 bf8_from(const float *v) {
    bf16x8 r;
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] = (__bf16)v[j];
    return r;
}

// B operand of K step s from four D tiles (tiles 2s and 2s+1 are used), after
// bias and ReLU have been applied in place.
__device__ __forceinline__ bf16x8 mlp_b_from_tiles(const f32x4 *t, int s) {
    bf16x8 r;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        r[j] = (__bf16)t[2 * s][j];
        r[j + 4] = (__bf16)t[2 * s + 1][j];
    }
    return r;
}

// Normalised input B operand for the point this lane's column stands for: lane
// group g holds features 8g .. 8g + 7 (zero beyond the 14th).
__device__ __forceinline__ bf16x8 mlp_input_operand(const MlpDev &m, const float *feat, int g) {
    float v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        int f = 8 * g + j;
        v[j] = (f < kMlpIn && feat) ? (feat[f] - m.norm[kNormInMean + f]) * m.norm[kNormInScale + f] : 0.0f;
    }
    return bf8_from(v);
}

struct MlpFwd {
    f32x4 z1[4], z2[4]; // pre-activations (D layout)
    f32x4 y;            // outputs 0 and 1 in elements 0, 1 of lanes 0-15
};

// Forward pass of 16 points (one per lane column); x is the input operand.
__device__ __forceinline__ void mlp_forward16(const MlpDev &m, bf16x8 x, MlpFwd &F) {
    const int lane = threadIdx.x & 63, g = lane >> 4;
    const bf16x8 *fr = m.frag + lane;
    const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
    f32x4 h[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        F.z1[t] = mfma(fr[64 * (kFragW1 + t)], x, zero);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            float z = F.z1[t][i] + m.params[kOffB1 + 16 * t + 4 * g + i];
            F.z1[t][i] = z;
            h[t][i] = z > 0.0f ? z : 0.0f;
        }
    }
    const bf16x8 b20 = mlp_b_from_tiles(h, 0), b21 = mlp_b_from_tiles(h, 1);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        f32x4 acc = mfma(fr[64 * (kFragW2 + 2 * t)], b20, zero);
        acc = mfma(fr[64 * (kFragW2 + 2 * t + 1)], b21, acc);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            float z = acc[i] + m.params[kOffB2 + 16 * t + 4 * g + i];
            F.z2[t][i] = z;
            h[t][i] = z > 0.0f ? z : 0.0f;
        }
    }
    const bf16x8 b30 = mlp_b_from_tiles(h, 0), b31 = mlp_b_from_tiles(h, 1);
    f32x4 y = mfma(fr[64 * kFragW3], b30, zero);
    y = mfma(fr[64 * (kFragW3 + 1)], b31, y);
    if (g == 0) {
        y[0] += m.params[kOffB3 + 0];
        y[1] += m.params[kOffB3 + 1];
    }
    F.y = y;
}

} // namespace dbx
