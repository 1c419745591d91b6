// learned.hip -- training of the per-list position models on the GPU.
//
// Input: the (docID, position) pairs of each modelled list as emitted by the
// decompression kernel (dec[dec_off[m] + i] = docID at position i).
//  * PLA-256: knot p holds the number of list elements whose scaled input t
//    lies below p -- the empirical CDF at 257 equally spaced knots, computed
//    by one binary search per knot (a closed-form least-error fit of a
//    monotone linear spline to the CDF at its knots).
//  * PLA-Q: knots at 256 equal position steps, x = the docID there as a
//    float, y = the position, and each piece's slope (read off the list).
//  * MLP 1-32-1 (ReLU): full-batch Adam on the mean squared position error
//    over `samples` evenly strided pairs; one workgroup per list, forward and
//    backward passes written out by hand, gradients reduced across the wave
//    with shuffles and across the workgroup through LDS.
//  * err: max |prediction - position| over every element of the list, with
//    the inference functions of exc_model.h -- the same code the search uses.
#include "exc_internal.h"
#include "exc_model.h"

namespace exc {
namespace {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

// input offset / scale of a model, computed identically by every thread
__device__ __forceinline__ void model_scale(const uint32_t* L, uint32_t n, uint32_t kind, float& x0, float& xs)
{
    x0 = (float)L[0];
    const float span = (float)L[n - 1] - x0 + 1.0f;
    xs = (kind == kModelPla) ? 256.0f / span : 1.0f / span;
}

__global__ __launch_bounds__(256) void k_pla_fit(TrainArgs a)
{
    const uint32_t m = blockIdx.x;
    Model& M = a.models[m];
    const uint32_t n = M.n;
    const uint32_t* L = a.dec + a.dec_off[m];
    float x0, xs;
    model_scale(L, n, kModelPla, x0, xs);
    float* y = a.params + M.param_off;
    for (uint32_t p = threadIdx.x; p < (uint32_t)kPlaKnots; p += blockDim.x) {
        // number of elements with ((float)v - x0) * xs < p
        uint32_t lo = 0, hi = n;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (((float)L[mid] - x0) * xs < (float)p) lo = mid + 1; else hi = mid;
        }
        y[p] = (float)lo;
    }
    if (threadIdx.x == 0) {
        M.x0 = x0;
        M.xs = xs;
    }
}

__global__ __launch_bounds__(256) void k_plaq_fit(TrainArgs a)
{
    const uint32_t m = blockIdx.x;
    Model& M = a.models[m];
    const uint32_t n = M.n;
    const uint32_t* L = a.dec + a.dec_off[m];
    float* p = a.params + M.param_off;
    const auto q = [&](uint32_t j) { return (uint32_t)(((uint64_t)j * (n - 1)) / 256u); };
    for (uint32_t j = threadIdx.x; j < (uint32_t)kPlaKnots; j += blockDim.x) {
        const float x = (float)L[q(j)], y = (float)q(j);
        p[j] = x;
        p[kPlaKnots + j] = y;
        if (j + 1 < (uint32_t)kPlaKnots) {
            const float dx = (float)L[q(j + 1)] - x;
            p[2 * kPlaKnots + j] = dx > 0.0f ? ((float)q(j + 1) - y) / dx : 0.0f;
        }
    }
    if (threadIdx.x == 0) {
        M.x0 = 0.0f;
        M.xs = 0.0f;
    }
}

// Deterministic initial parameters: hinges at j/32 with unit input weight,
// output = input (the CDF of a uniform list), plus a small seeded jitter.
__device__ __forceinline__ float mlp_init(uint32_t j, uint64_t seed)
{
    uint64_t z = seed * 0x9E3779B97F4A7C15ull + j * 0xBF58476D1CE4E5B9ull + 1;
    z = (z ^ (z >> 31)) * 0x94D049BB133111EBull;
    const float jit = ((float)((z >> 40) & 0xffff) / 65536.0f - 0.5f) * 1.0e-3f;
    if (j < (uint32_t)kMlpH) return 1.0f;                                  // w1
    if (j < 2u * kMlpH) return -(float)(j - kMlpH) / (float)kMlpH;        // b1
    if (j < 3u * kMlpH) return (j == 2u * kMlpH ? 1.0f : 0.0f) + jit;     // w2
    return 0.0f;                                                           // b2
}

__global__ __launch_bounds__(256) void k_mlp_train(TrainArgs a)
{
    __shared__ float w[kMlpParams];
    __shared__ float part[4][kMlpParams + 1];
    const uint32_t m = blockIdx.x, tid = threadIdx.x, ln = lane_id(), wid = tid >> 6;
    Model& M = a.models[m];
    const uint32_t n = M.n;
    const uint32_t* L = a.dec + a.dec_off[m];
    float x0, xs;
    model_scale(L, n, kModelMlp, x0, xs);
    float* out = a.params + M.param_off;
    const uint32_t S = a.samples < n ? a.samples : n;
    float mom = 0.0f, vel = 0.0f;
    if (tid < (uint32_t)kMlpParams) w[tid] = (a.epochs == 0) ? out[tid] : mlp_init(tid, a.seed + m);
    __syncthreads();
    const uint32_t iters = a.epochs == 0 ? 1 : a.epochs;
    for (uint32_t ep = 0; ep < iters; ++ep) {
        float g1[kMlpH], gb1[kMlpH], g2[kMlpH], gb2 = 0.0f, loss = 0.0f;
#pragma unroll
        for (int j = 0; j < kMlpH; ++j) { g1[j] = 0.0f; gb1[j] = 0.0f; g2[j] = 0.0f; }
        for (uint32_t s = tid; s < S; s += blockDim.x) {
            const uint32_t i = (uint32_t)(((uint64_t)s * n) / S);
            const float x = ((float)L[i] - x0) * xs;
            const float t = (float)i / (float)n;
            float h[kMlpH], o = w[3 * kMlpH];
#pragma unroll
            for (int j = 0; j < kMlpH; ++j) {
                h[j] = fmaxf(fmaf(w[j], x, w[kMlpH + j]), 0.0f);
                o = fmaf(w[2 * kMlpH + j], h[j], o);
            }
            const float e = o - t;
            loss = fmaf(e, e, loss);
            const float g = 2.0f * e / (float)S;  // d(mean squared error)/d(output)
            gb2 += g;
#pragma unroll
            for (int j = 0; j < kMlpH; ++j) {
                g2[j] = fmaf(g, h[j], g2[j]);
                const float d = h[j] > 0.0f ? g * w[2 * kMlpH + j] : 0.0f;
                g1[j] = fmaf(d, x, g1[j]);
                gb1[j] += d;
            }
        }
        // reduce the 97 gradients (and the loss) over the workgroup
#pragma unroll
        for (int j = 0; j < kMlpH; ++j) {
            const float r1 = wave_sum(g1[j]), rb = wave_sum(gb1[j]), r2 = wave_sum(g2[j]);
            if (ln == 0) { part[wid][j] = r1; part[wid][kMlpH + j] = rb; part[wid][2 * kMlpH + j] = r2; }
        }
        {
            const float rb2 = wave_sum(gb2), rl = wave_sum(loss);
            if (ln == 0) { part[wid][3 * kMlpH] = rb2; part[wid][kMlpParams] = rl; }
        }
        __syncthreads();
        if (tid <= (uint32_t)kMlpParams) {
            const float g = part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
            if (a.epochs == 0) {
                a.grad_out[m * (kMlpParams + 1) + tid] = (tid == (uint32_t)kMlpParams) ? g / (float)S : g;
            } else if (tid < (uint32_t)kMlpParams) {
                // Adam (beta1 0.9, beta2 0.999, eps 1e-12)
                mom = 0.9f * mom + 0.1f * g;
                vel = 0.999f * vel + 0.001f * g * g;
                const float mh = mom / (1.0f - __powf(0.9f, (float)(ep + 1)));
                const float vh = vel / (1.0f - __powf(0.999f, (float)(ep + 1)));
                w[tid] -= a.lr * mh / (sqrtf(vh) + 1.0e-12f);
            }
        }
        __syncthreads();
    }
    if (a.epochs != 0 && tid < (uint32_t)kMlpParams) out[tid] = w[tid];
    if (tid == 0 && a.epochs != 0) {
        M.x0 = x0;
        M.xs = xs;
    }
}

// max |pred - position| per model; grid.y = model, grid.x strides the list
__global__ __launch_bounds__(256) void k_model_err(TrainArgs a)
{
    const uint32_t m = blockIdx.y;
    const Model M = a.models[m];
    const uint32_t* L = a.dec + a.dec_off[m];
    const float* prm = a.params + M.param_off;
    float e = 0.0f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < M.n; i += gridDim.x * blockDim.x) {
        const float p = model_pred(M.kind, prm, M.x0, M.xs, M.n, L[i]);
        e = fmaxf(e, fabsf(p - (float)i));
    }
    e = wave_max(e);
    if (lane_id() == 0 && e > 0.0f) atomicMax(a.err_bits + m, __float_as_uint(e));
}

}  // namespace

hipError_t launch_pla_fit(const TrainArgs& a, hipStream_t s)
{
    if (a.nmodels) hipLaunchKernelGGL(k_pla_fit, dim3(a.nmodels), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_plaq_fit(const TrainArgs& a, hipStream_t s)
{
    if (a.nmodels) hipLaunchKernelGGL(k_plaq_fit, dim3(a.nmodels), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_mlp_train(const TrainArgs& a, hipStream_t s)
{
    if (a.nmodels) hipLaunchKernelGGL(k_mlp_train, dim3(a.nmodels), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_model_err(const TrainArgs& a, uint64_t total, hipStream_t s)
{
    if (!a.nmodels) return hipSuccess;
    uint64_t per = (total / a.nmodels + 255) / 256;
    const uint32_t gx = (uint32_t)(per < 1 ? 1 : (per > 1024 ? 1024 : per));
    hipLaunchKernelGGL(k_model_err, dim3(gx, a.nmodels), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace exc
