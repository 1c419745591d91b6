// exc_model.h -- fp32 inference of the learned position models, shared by the
// training/error kernels (learned.hip) and the probe kernel (kernels.hip).
//
// A model maps a docID v of one list to its (fractional) position in the list:
//   PLA  t = clamp((v - x0) * xs, 0, 256), p = min(int(t), 255),
//        pos = y[p] + (t - p) * (y[p+1] - y[p])      (257 knot values y)
//   MLP  x = (v - x0) * xs, h_j = max(0, w1_j x + b1_j), j < 32,
//        pos = n * (b2 + sum_j w2_j h_j)             (params w1[32] b1[32] w2[32] b2)
//   PLA-Q knots at equal position steps: q_j = floor(j (n - 1) / 256),
//        x_j = (float)docID at q_j, y_j = q_j, s_j = (y_(j+1) - y_j) / (x_(j+1) - x_j)
//        (0 where the knots coincide); piece p = the last knot with x_p <= v
//        (8 halvings over the 257 knots), pos = y_p + ((float)v - x_p) * s_p,
//        y_0 below x_0 and y_256 above x_256      (params x[257] y[257] s[256])
// Every multiply-add is an explicit fmaf so the oracle can repeat the same
// arithmetic bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace exc {

// The list-level line of a d-gap (ParaPFD) list, derived from its block heads
// at upload (the reference's ParaPFD format carries no list fit): docID v sits
// near position (v - icpt) / slope, slope = docIDs per position between the
// first and last head, and every head lies within [-rr, +rl] positions of the
// line, widened by one block.  The probe's head window uses it as it uses the
// SegLRC list fit, and verifies the window against the heads, so the line only
// has to be close.  Lists of fewer than two blocks (or a flat line) get the
// no-window values.  Host and device compute it with the same float steps.
__host__ __device__ inline void heads_line(const uint32_t* hd, uint32_t nseg, uint32_t seg, float& slope, float& icpt,
                                           float& rl, float& rr)
{
    slope = 1.0f;
    icpt = 0.0f;
    rl = rr = 4.0e9f;
    if (nseg < 2 || hd[nseg - 1] <= hd[0]) return;
    const float s = (float)((double)(hd[nseg - 1] - hd[0]) / ((double)(nseg - 1) * (double)seg));
    if (!(s > 0.0f)) return;
    const float c = (float)hd[0], inv = 1.0f / s;
    float dl = 0.0f, dr = 0.0f;
    for (uint32_t b = 0; b < nseg; ++b) {
        const float d = ((float)hd[b] - c) * inv - (float)((double)b * (double)seg);
        dl = d > dl ? d : dl;
        dr = -d > dr ? -d : dr;
    }
    slope = s;
    icpt = c;
    rl = dl + (float)seg;
    rr = dr + (float)seg;
}

constexpr uint32_t kModelPla = 1;
constexpr uint32_t kModelMlp = 2;
constexpr uint32_t kModelPlaQ = 3;
constexpr int kPlaKnots = 257;
constexpr int kMlpH = 32;
constexpr int kMlpParams = 3 * kMlpH + 1;  // w1, b1, w2, b2
constexpr int kPlaQParams = 3 * kPlaKnots - 1;  // x, y, s
// floats per model slot in the parameter pool (padded)
constexpr uint32_t kPlaStride = 260;
constexpr uint32_t kMlpStride = 100;
constexpr uint32_t kPlaQStride = 772;
__host__ __device__ __forceinline__ uint32_t model_stride(uint32_t kind)
{
    return kind == kModelPla ? kPlaStride : kind == kModelPlaQ ? kPlaQStride : kMlpStride;
}
__host__ __device__ __forceinline__ uint32_t model_nparams(uint32_t kind)
{
    return kind == kModelPla ? (uint32_t)kPlaKnots : kind == kModelPlaQ ? (uint32_t)kPlaQParams : (uint32_t)kMlpParams;
}

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

__host__ __device__ __forceinline__ float plaq_pred(const float* p, uint32_t v)
{
    const float* x = p;
    const float* y = p + kPlaKnots;
    const float* sl = p + 2 * kPlaKnots;
    const float fv = (float)v;
    if (fv < x[0]) return y[0];
    if (fv >= x[kPlaKnots - 1]) return y[kPlaKnots - 1];
    int lo = 0, hi = kPlaKnots - 1;  // x[lo] <= fv < x[hi]
    while (hi - lo > 1) {
        const int m = (lo + hi) >> 1;
        if (x[m] <= fv) lo = m; else hi = m;
    }
    return fmaf(fv - x[lo], sl[lo], y[lo]);
}

// prediction of any model kind
__host__ __device__ __forceinline__ float model_pred(uint32_t kind, const float* prm, float x0, float xs, uint32_t n,
                                                     uint32_t v)
{
    if (kind == kModelPla) return pla_pred(prm, x0, xs, v);
    if (kind == kModelPlaQ) return plaq_pred(prm, v);
    return mlp_pred(prm, x0, xs, n, v);
}

}  // namespace exc
