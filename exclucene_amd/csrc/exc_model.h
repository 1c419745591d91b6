// exc_model.h -- fp32 inference of the learned position models, shared by the
// training/error kernels (learned.hip) and the probe kernel (kernels.hip).
//
// A model maps a docID v of one list to its (fractional) position in the list:
//   PLA  t = clamp((v - x0) * xs, 0, 256), p = min(int(t), 255),
//        pos = y[p] + (t - p) * (y[p+1] - y[p])      (257 knot values y)
//   MLP  x = (v - x0) * xs, h_j = max(0, w1_j x + b1_j), j < 32,
//        pos = n * (b2 + sum_j w2_j h_j)             (params w1[32] b1[32] w2[32] b2)
// Every multiply-add is an explicit fmaf so the oracle can repeat the same
// arithmetic bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace exc {

constexpr uint32_t kModelPla = 1;
constexpr uint32_t kModelMlp = 2;
constexpr int kPlaKnots = 257;
constexpr int kMlpH = 32;
constexpr int kMlpParams = 3 * kMlpH + 1;  // w1, b1, w2, b2
// floats per model slot in the parameter pool (padded)
constexpr uint32_t kPlaStride = 260;
constexpr uint32_t kMlpStride = 100;
__host__ __device__ __forceinline__ uint32_t model_stride(uint32_t kind) { return kind == kModelPla ? kPlaStride : kMlpStride; }

__host__ __device__ __forceinline__ float pla_pred(const float* y, float x0, float xs, uint32_t v)
{
    float t = ((float)v - x0) * xs;
    t = fminf(fmaxf(t, 0.0f), 256.0f);
    int p = (int)t;
    p = p > 255 ? 255 : p;
    const float f = t - (float)p;
    return fmaf(f, y[p + 1] - y[p], y[p]);
}

__host__ __device__ __forceinline__ float mlp_pred(const float* w, float x0, float xs, uint32_t n, uint32_t v)
{
    const float x = ((float)v - x0) * xs;
    float acc = w[3 * kMlpH];
    for (int j = 0; j < kMlpH; ++j) {
        const float h = fmaxf(fmaf(w[j], x, w[kMlpH + j]), 0.0f);
        acc = fmaf(w[2 * kMlpH + j], h, acc);
    }
    return acc * (float)n;
}

}  // namespace exc
