// exc_encode.h -- the SegLRC block fit and residuals, shared by the host index
// builder (builder.cpp) and the device index builder (gpubuild.hip) so both
// produce the same bytes.  Both sides are compiled with -ffp-contract=off, so
// every double and float operation below rounds the same way on the CPU and
// on gfx950.
//
// Reference: the per-segment least-squares fit and ranges of
// GC/SegLRC/SegLRC_Generator/SegLRC_Generator.cu:264-331 (fit in double, slope
// and intercept stored as float, nRange = (uint)max deviation + 1) and the
// residual base of GC/SegLRC/SegLRC_Generator/SegLRC_Generator.cu:185-233.
#pragma once
#include "../../include/exclucene_amd.h"
#include "exc_format.h"

namespace exc {

EXC_HD exc_seg_info seg_fit(const uint32_t* L, uint32_t n)
{
    exc_seg_info r{1.0f, 0.0f, 0, 0};
    if (n == 0) return r;
    if (n == 1) {
        r.fIntercept = (float)L[0];
        return r;
    }
    double xa = 0, ya = 0;
    for (uint32_t i = 0; i < n; ++i) {
        xa += i;
        ya += L[i];
    }
    xa /= n;
    ya /= n;
    double sx = 0, sxy = 0;
    for (uint32_t i = 0; i < n; ++i) {
        double dx = i - xa;
        sx += dx * dx;
        sxy += dx * ((double)L[i] - ya);
    }
    r.fSlope = (float)(sxy / sx);
    r.fIntercept = (float)(ya - (double)r.fSlope * xa);
    float rl = 0, rr = 0;
    for (uint32_t i = 0; i < n; ++i) {
        float d = ((float)L[i] - r.fIntercept) / r.fSlope - (float)i;
        if (d > rl) rl = d;
        else if (d < rr) rr = d;
    }
    r.nRangeLeft = (uint32_t)rl + 1;
    r.nRangeRight = (uint32_t)(0 - rr) + 1;
    return r;
}

// Residual of element j against a segment fit (before the base shift).
EXC_HD int32_t seg_residual(const uint32_t* L, uint32_t j, const exc_seg_info& f)
{
    return (int32_t)(L[j] - (uint32_t)seg_y(j, f.fSlope, f.fIntercept));
}

// Base (minimum residual) and the bit width of the largest shifted residual.
EXC_HD void seg_base_bits(const uint32_t* L, uint32_t n, const exc_seg_info& f, int32_t& base, uint32_t& lb)
{
    int32_t mn = 0, mx = 0;
    for (uint32_t j = 0; j < n; ++j) {
        const int32_t d = seg_residual(L, j, f);
        if (j == 0 || d < mn) mn = d;
        if (j == 0 || d > mx) mx = d;
    }
    base = n ? mn : 0;
    lb = bits_of((uint32_t)(mx - mn));
}

}  // namespace exc
