// exc_format.h -- layout constants and the pinned reconstruction formulas shared
// by the index builder (host) and the gfx950 kernels (device).
//
// Reference formulas (GC = src/gpu/Compressed/src):
//   LRC    y(i) = int((float)i * fSlope) + iIntercept, i list-global
//          (GC/LRC/LRC_Intersection/LRC_Intersection_kernel.cu:255,344)
//   SegLRC y(j) = (int)(j * fSlope + fIntercept) + base, j in-segment
//          (GC/SegLRC/SegLRC_Intersection/SegLRC_Intersection_kernel.cu:266,382);
//          nvcc contracts j*s+b to one fused multiply-add, which is pinned here
//          as an explicit fmaf on both sides.
//   ParaPFor descriptor lb<<26 | eb<<21 | hb<<16 | en (GI/ParaPFor.cpp:88-91).
//   NewPFor descriptor b<<26 | en, then the b-bit low parts, then the
//          exceptions as one Simple16 stream [pos0, pos_i - pos_(i-1) - 1 ...,
//          high parts ...] (GI/NewPFor.cpp:54-99); lb and en sit where the
//          ParaPFor descriptor keeps them.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define EXC_HD __host__ __device__ __forceinline__

namespace exc {

// Zero words after every device array read as value pairs.  Block decoders
// read field j of a block for every j < the segment size (256 at most)
// without clamping j to the block's length: the highest word such a read
// touches is word (255 * 31) / 32 + 1 = 248 of the block's fields, so 256
// words of padding keep the reads of an array's last block inside it.
constexpr uint32_t kPadWords = 256;
constexpr uint32_t kMaxSeg = 256;
static_assert(((kMaxSeg - 1) * 31) / 32 + 2 <= kPadWords, "unclamped field reads stay inside the padding");

// float -> int32 as the reference's GPU code gets it (CUDA's cvt.rzi.s32.f32)
// and as gfx950's v_cvt_i32_f32 does it: toward zero, out-of-range values
// saturated, NaN to 0.  A C cast is undefined out of range (x86 gives
// INT32_MIN), so encoder and decoder both go through this: on the device as
// the one instruction (the range checks written out cost the probe 8 %),
// on the host written out.
EXC_HD int32_t f2i_sat(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    int32_t r;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
#else
    if (!(x == x)) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x < -2147483648.0f) return (int32_t)0x80000000u;
    return (int32_t)x;
#endif
}

EXC_HD int32_t lrc_y(uint32_t i, float slope) { return f2i_sat((float)i * slope); }
EXC_HD int32_t seg_y(uint32_t j, float slope, float icpt) { return f2i_sat(fmaf((float)j, slope, icpt)); }
// the predictions the residuals are taken against, added modulo 2^32
EXC_HD uint32_t lrc_pred(uint32_t i, float slope, int32_t icpt) { return (uint32_t)lrc_y(i, slope) + (uint32_t)icpt; }
EXC_HD uint32_t seg_pred(uint32_t j, float slope, float icpt, int32_t base)
{
    return (uint32_t)seg_y(j, slope, icpt) + (uint32_t)base;
}

EXC_HD uint32_t desc_lb(uint32_t d) { return (d >> 26) & 31u; }
EXC_HD uint32_t desc_eb(uint32_t d) { return (d >> 21) & 31u; }
EXC_HD uint32_t desc_hb(uint32_t d) { return (d >> 16) & 31u; }
EXC_HD uint32_t desc_en(uint32_t d) { return d & 65535u; }
EXC_HD uint32_t words_for(uint32_t n, uint32_t b) { return (uint32_t)(((uint64_t)n * b + 31) >> 5); }
// Quarter boundaries of a ParaPFD block of len values: m1, m2 = len / 2 (the
// .segMedian position) and m3; the device records carry the values at m1 and
// m3 besides the median, so a probe decodes at most a quarter of a block.
EXC_HD uint32_t pfd_m1(uint32_t len) { return len >> 2; }
EXC_HD uint32_t pfd_m3(uint32_t len) { return (len >> 1) + ((len - (len >> 1)) >> 1); }
// Start of quarter q (0..4, 4 = len) and the midpoint of quarter q (0..3):
// the eighth boundaries whose values a device index keeps per block.
EXC_HD uint32_t pfd_qb(uint32_t len, uint32_t q)
{
    return q == 0 ? 0u : q == 1 ? pfd_m1(len) : q == 2 ? (len >> 1) : q == 3 ? pfd_m3(len) : len;
}
EXC_HD uint32_t pfd_oct_pos(uint32_t len, uint32_t q)
{
    return pfd_qb(len, q) + ((pfd_qb(len, q + 1) - pfd_qb(len, q)) >> 1);
}
EXC_HD uint32_t bits_of(uint32_t v) { return v ? 32u - (uint32_t)__builtin_clz(v) : 0u; }

// Simple16 (Zhang, Long and Suel, WWW 2008; the tables of GI/Simple16.h:13-48):
// a word is a 4-bit selector k in bits 28-31 and fields packed from bit 0 up;
// layout k is up to three runs of equal-width fields.  Packed per selector:
// five bits each of (count, width) for runs 0, 1, 2.
constexpr uint32_t s16_runs(uint32_t c0, uint32_t w0, uint32_t c1 = 0, uint32_t w1 = 0, uint32_t c2 = 0,
                            uint32_t w2 = 0)
{
    return c0 | w0 << 5 | c1 << 10 | w1 << 15 | c2 << 20 | w2 << 25;
}

EXC_HD constexpr uint32_t s16_layout(uint32_t k)
{
    switch (k & 15u) {
    case 0: return s16_runs(28, 1);
    case 1: return s16_runs(7, 2, 14, 1);
    case 2: return s16_runs(7, 1, 7, 2, 7, 1);
    case 3: return s16_runs(14, 1, 7, 2);
    case 4: return s16_runs(14, 2);
    case 5: return s16_runs(1, 4, 8, 3);
    case 6: return s16_runs(1, 3, 4, 4, 3, 3);
    case 7: return s16_runs(7, 4);
    case 8: return s16_runs(4, 5, 2, 4);
    case 9: return s16_runs(2, 4, 4, 5);
    case 10: return s16_runs(3, 6, 2, 5);
    case 11: return s16_runs(2, 5, 3, 6);
    case 12: return s16_runs(4, 7);
    case 13: return s16_runs(1, 10, 2, 9);
    case 14: return s16_runs(2, 14);
    default: return s16_runs(1, 28);
    }
}

// number of fields of a Simple16 word of layout L (all are decoded, used or not)
EXC_HD uint32_t s16_count_l(uint32_t L) { return (L & 31u) + ((L >> 10) & 31u) + ((L >> 20) & 31u); }
EXC_HD uint32_t s16_count(uint32_t word) { return s16_count_l(s16_layout(word >> 28)); }

// field k of a Simple16 word of layout L, k < its count (larger k: some
// value, never a shift of 32 or more)
EXC_HD uint32_t s16_field_l(uint32_t word, uint32_t k, uint32_t L)
{
    const uint32_t c0 = L & 31u, w0 = (L >> 5) & 31u, c1 = (L >> 10) & 31u, w1 = (L >> 15) & 31u;
    const uint32_t w2 = (L >> 25) & 31u;
    uint32_t sh, w;
    if (k < c0) {
        sh = k * w0;
        w = w0;
    } else if (k < c0 + c1) {
        sh = c0 * w0 + (k - c0) * w1;
        w = w1;
    } else {
        sh = c0 * w0 + c1 * w1 + (k - c0 - c1) * w2;
        w = w2;
    }
    return (word >> (sh < 31u ? sh : 31u)) & ((1u << w) - 1u);
}
EXC_HD uint32_t s16_field(uint32_t word, uint32_t k) { return s16_field_l(word, k, s16_layout(word >> 28)); }

// the fields of one Simple16 word into out[0 .. min(count, lim))
EXC_HD void s16_unpack(uint32_t word, uint32_t* out, uint32_t lim)
{
    const uint32_t L = s16_layout(word >> 28);
    uint32_t sh = 0, o = 0;
    for (uint32_t r = 0; r < 3; ++r) {
        const uint32_t c = (L >> (10 * r)) & 31u, w = (L >> (10 * r + 5)) & 31u;
        for (uint32_t i = 0; i < c; ++i, ++o, sh += w)
            if (o < lim) out[o] = (word >> sh) & ((1u << w) - 1u);
    }
}

}  // namespace exc
