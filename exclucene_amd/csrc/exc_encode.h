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

// LRC list fit, GC/LRC/LRC_Generator/LRC_Generator.cu:221-287: double least
// squares over (i, docID), slope stored as float, intercept truncated to int,
// search range = the spread of the deviations (position units).
EXC_HD exc_lrc_info lrc_fit(const uint32_t* L, uint32_t n)
{
    exc_lrc_info r{1.0f, 0, 0.0f};
    if (n == 0) return r;
    if (n == 1) {
        r.iIntercept = (int32_t)L[0];
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
    double slope = sxy / sx, icpt = ya - slope * xa, rl = 0, rr = 0;
    for (uint32_t i = 0; i < n; ++i) {
        double d = (L[i] - icpt) / slope - i;
        if (d > rl) rl = d;
        else if (d < rr) rr = d;
    }
    r.fSlope = (float)slope;
    r.iIntercept = (int32_t)icpt;
    r.fSearchRange = (float)(rl - rr);
    return r;
}

// The LRC residuals' minimum, which the generator folds into the intercept so
// that every residual is >= 0 (GC/LRC/LRC_Generator/LRC_Generator.cu:165-180).
EXC_HD int32_t lrc_min_residual(const uint32_t* L, uint32_t n, const exc_lrc_info& f)
{
    int32_t mn = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const int32_t d = (int32_t)(L[i] - lrc_pred(i, f.fSlope, f.iIntercept));
        if (i == 0 || d < mn) mn = d;
    }
    return n > 1 ? mn : 0;
}

// ---- ParaPFor blocks (GI/ParaPFor.cpp:31-94) over values val(0 .. n), n <= kMaxSeg

// Smallest b with #(v >= 2^b) <= frac * n, compared in float (GI/ParaPFor.cpp:31-53).
template <class F>
EXC_HD uint32_t pfor_best_b(F val, uint32_t n, float frac)
{
    uint32_t hist[33] = {0};
    for (uint32_t i = 0; i < n; ++i) hist[bits_of(val(i))]++;
    uint32_t above = n - hist[0];  // values needing more than b bits, b = 0
    for (uint32_t b = 0; b < 32; ++b) {
        if ((float)above <= frac * (float)n) return b;
        above -= hist[b + 1];
    }
    return 32;
}

// Descriptor and word count (descriptor included) of one block; words = 0
// when the descriptor cannot hold it (lb or a high part of 32 bits, more than
// 65,535 exceptions).
struct PforPlan {
    uint32_t desc = 0, words = 0;
};

template <class F>
EXC_HD PforPlan pfor_plan(F val, uint32_t n, float frac)
{
    PforPlan p;
    const uint32_t lb = pfor_best_b(val, n, frac);
    if (lb > 31) return p;
    uint32_t en = 0, last = 0, hb = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t x = val(i);
        if (bits_of(x) > lb) {
            ++en;
            last = i;
            const uint32_t h = bits_of(x >> lb);
            hb = h > hb ? h : hb;
        }
    }
    const uint32_t eb = en ? bits_of(last) : 0;
    if (en > 65535 || hb > 31) return p;
    p.desc = en ? (lb << 26) | (eb << 21) | (hb << 16) | en : (lb << 26);
    p.words = 1 + words_for(n, lb) + words_for(en, eb) + words_for(en, hb);
    return p;
}

// LSB-first bit packing into consecutive words (GI/pack.h:14-24), each word
// written once.
struct BitWriter {
    uint32_t* o;
    uint64_t acc = 0;
    uint32_t nb = 0;
    EXC_HD explicit BitWriter(uint32_t* out) : o(out) {}
    EXC_HD void put(uint32_t v, uint32_t b)
    {
        if (!b) return;
        acc |= (uint64_t)v << nb;
        nb += b;
        if (nb >= 32) {
            *o++ = (uint32_t)acc;
            acc >>= 32;
            nb -= 32;
        }
    }
    EXC_HD void flush()  // to the next word boundary
    {
        if (nb) *o++ = (uint32_t)acc;
        acc = 0;
        nb = 0;
    }
};

// The block's words: descriptor, lb-bit low parts, then the exceptions'
// positions (eb bits) and high parts (hb bits), each section from a word
// boundary (GI/ParaPFor.cpp:56-94).
template <class F>
EXC_HD void pfor_write(F val, uint32_t n, uint32_t desc, uint32_t* out)
{
    const uint32_t lb = desc_lb(desc), eb = desc_eb(desc), hb = desc_hb(desc), en = desc_en(desc);
    const uint32_t mask = (1u << lb) - 1u;
    out[0] = desc;
    BitWriter w(out + 1);
    for (uint32_t i = 0; i < n; ++i) w.put(val(i) & mask, lb);
    w.flush();
    if (!en) return;
    for (uint32_t i = 0; i < n; ++i)
        if (bits_of(val(i)) > lb) w.put(i, eb);
    w.flush();
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t x = val(i);
        if (bits_of(x) > lb) w.put(x >> lb, hb);
    }
    w.flush();
}

}  // namespace exc
