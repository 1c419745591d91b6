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
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define EXC_HD __host__ __device__ __forceinline__

namespace exc {

constexpr uint32_t kPadWords = 64;   // zero words after every device array read as value pairs

EXC_HD int32_t lrc_y(uint32_t i, float slope) { return (int32_t)((float)i * slope); }
EXC_HD int32_t seg_y(uint32_t j, float slope, float icpt) { return (int32_t)fmaf((float)j, slope, icpt); }

EXC_HD uint32_t desc_lb(uint32_t d) { return (d >> 26) & 31u; }
EXC_HD uint32_t desc_eb(uint32_t d) { return (d >> 21) & 31u; }
EXC_HD uint32_t desc_hb(uint32_t d) { return (d >> 16) & 31u; }
EXC_HD uint32_t desc_en(uint32_t d) { return d & 65535u; }
EXC_HD uint32_t words_for(uint32_t n, uint32_t b) { return (uint32_t)(((uint64_t)n * b + 31) >> 5); }
EXC_HD uint32_t bits_of(uint32_t v) { return v ? 32u - (uint32_t)__builtin_clz(v) : 0u; }

}  // namespace exc
