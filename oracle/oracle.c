/*
 * oracle.c -- CPU restatement of ExCLucene's GPU decode + intersect path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): used by tests/, smoke() and the
 * cpu_baseline leg of bench.py.  The exclucene_amd package never links it.
 * Built with -ffp-contract=off so that every fp32 operation is the one written.
 * Each function names the reference file:line it restates (paths relative to
 * the reference root; GI = src/gpu/Compressed/include, GC = src/gpu/Compressed/src).
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* GI/util.h:13-27 */
static inline uint64_t div_roundup(uint64_t v, uint32_t d) { return (v + (d - 1)) / d; }
static inline uint32_t gccbits(uint32_t v) { return v == 0 ? 0 : 32 - (uint32_t)__builtin_clz(v); }
static uint32_t maxbits(const uint32_t *in, uint32_t n)
{
    uint32_t acc = 0;
    for (uint32_t i = 0; i < n; ++i) acc |= in[i];
    return gccbits(acc);
}

/* GI/pack.h:14-24: n values of b bits, LSB-first, a value may straddle two words. */
static void pack_bits(const uint32_t *in, uint32_t n, uint32_t b, uint32_t *out)
{
    uint32_t nw = (uint32_t)div_roundup((uint64_t)n * b, 32);
    memset(out, 0, (size_t)nw * 4);
    if (b == 0) return;
    for (uint32_t i = 0; i < n; ++i) {
        uint64_t bit = (uint64_t)i * b;
        uint32_t idx = (uint32_t)(bit >> 5), sh = (uint32_t)(bit & 31);
        out[idx] |= in[i] << sh;
        if (sh + b > 32) out[idx + 1] |= in[i] >> (32 - sh);
    }
}

/* GI/unpack.h:14-23 (reads the second word only when the value straddles). */
static void unpack_bits(const uint32_t *in, uint32_t n, uint32_t b, uint32_t *out)
{
    if (b == 0) { memset(out, 0, (size_t)n * 4); return; }
    uint32_t mask = (b == 32) ? 0xffffffffu : ((1u << b) - 1);
    for (uint32_t i = 0; i < n; ++i) {
        uint64_t bit = (uint64_t)i * b;
        uint32_t idx = (uint32_t)(bit >> 5), sh = (uint32_t)(bit & 31);
        uint64_t cw = in[idx];
        if (sh + b > 32) cw |= (uint64_t)in[idx + 1] << 32;
        out[i] = (uint32_t)(cw >> sh) & mask;
    }
}

/* GI/ParaPFor.cpp:31-53: smallest b whose exception count is <= FRAC * n (in float). */
uint32_t orc_pfor_best_b(const uint32_t *in, uint32_t n, float frac)
{
    for (uint32_t b = 0; b < 32; ++b) {
        uint32_t ex = 0;
        for (uint32_t i = 0; i < n; ++i)
            if (in[i] >= (1u << b)) ++ex;
        if ((float)ex <= frac * (float)n) return b;
    }
    return 32;
}

/* GI/ParaPFor.cpp:56-94. Descriptor = lb<<26 | eb<<21 | hb<<16 | en. */
uint32_t orc_pfor_encode(const uint32_t *in, uint32_t n, float frac, uint32_t *out)
{
    uint32_t lb = orc_pfor_best_b(in, n, frac);
    if (lb >= 32) return 0; /* lb = 32 does not fit the 5-bit field (ParaPFor.cpp:88,103) */
    uint32_t *lower = (uint32_t *)malloc((size_t)(n ? n : 1) * 4);
    uint32_t *hi = (uint32_t *)malloc((size_t)(n ? n : 1) * 4);
    uint32_t *pos = (uint32_t *)malloc((size_t)(n ? n : 1) * 4);
    uint32_t en = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (in[i] >= (1u << lb)) {
            lower[i] = in[i] & ((1u << lb) - 1);
            hi[en] = in[i] >> lb;
            pos[en] = i;
            ++en;
        } else {
            lower[i] = in[i];
        }
    }
    uint32_t w = 1;
    pack_bits(lower, n, lb, out + w);
    w += (uint32_t)div_roundup((uint64_t)n * lb, 32);
    uint32_t ret = 0;
    if (en > 0) {
        uint32_t eb = gccbits(pos[en - 1]);
        uint32_t hb = maxbits(hi, en);
        if (en > 65535 || hb > 31) goto done; /* unrepresentable descriptor */
        pack_bits(pos, en, eb, out + w);
        w += (uint32_t)div_roundup((uint64_t)en * eb, 32);
        pack_bits(hi, en, hb, out + w);
        w += (uint32_t)div_roundup((uint64_t)en * hb, 32);
        out[0] = (lb << 26) + (eb << 21) + (hb << 16) + en;
    } else {
        out[0] = lb << 26;
    }
    ret = w;
done:
    free(lower); free(hi); free(pos);
    return ret;
}

/* GI/ParaPFor.cpp:101-126 */
uint32_t orc_pfor_decode(const uint32_t *in, uint32_t n, uint32_t *out)
{
    uint32_t d = in[0];
    uint32_t lb = (d >> 26) % 32, eb = (d >> 21) & 31, hb = (d >> 16) & 31, en = d & 65535;
    uint32_t w = 1;
    unpack_bits(in + w, n, lb, out);
    w += (uint32_t)div_roundup((uint64_t)n * lb, 32);
    if (en > 0) {
        uint32_t *pos = (uint32_t *)malloc((size_t)en * 4), *hi = (uint32_t *)malloc((size_t)en * 4);
        unpack_bits(in + w, en, eb, pos);
        w += (uint32_t)div_roundup((uint64_t)en * eb, 32);
        unpack_bits(in + w, en, hb, hi);
        w += (uint32_t)div_roundup((uint64_t)en * hb, 32);
        for (uint32_t i = 0; i < en; ++i)
            if (pos[i] < n) out[pos[i]] |= hi[i] << lb;
        free(pos); free(hi);
    }
    return w;
}

/* ------------------------------------------------------------------ LRC */

/* GC/LRC/LRC_Generator/LRC_Generator.cu:221-287 (double-precision least squares). */
void orc_lrc_fit(const uint32_t *L, uint32_t n, orc_lrc_info *info)
{
    if (n == 0) { info->fSlope = 1; info->iIntercept = 0; info->fSearchRange = 0; return; }
    if (n == 1) { info->fSlope = 1; info->iIntercept = (int32_t)L[0]; info->fSearchRange = 0; return; }
    double xa = 0, ya = 0;
    for (uint32_t i = 0; i < n; ++i) { xa += i; ya += L[i]; }
    xa /= n; ya /= n;
    double sx = 0, sxy = 0, sy = 0;
    for (uint32_t i = 0; i < n; ++i) {
        double y = (double)L[i];
        sx += (double)((i - xa) * (i - xa));
        sxy += (double)((i - xa) * (y - ya));
        sy += (double)((y - ya) * (y - ya));
    }
    double slope = sxy / sx;
    double icpt = ya - slope * xa;
    double rl = 0, rr = 0;
    for (uint32_t i = 0; i < n; ++i) {
        double px = (L[i] - icpt) / slope;
        if (px - i > rl) rl = px - i;
        else if (px - i < rr) rr = px - i;
    }
    info->fSlope = (float)slope;
    info->iIntercept = (int32_t)icpt;
    info->fSearchRange = (float)(rl - rr);
}

/* float -> int32 as CUDA's cvt.rzi.s32.f32 (the reference's generator and
 * decoders run on the GPU): toward zero, saturating, NaN -> 0.  A C cast is
 * undefined out of range (x86 returns INT32_MIN). */
static inline int32_t f2i_sat(float x)
{
    if (x != x) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x < -2147483648.0f) return (int32_t)0x80000000u;
    return (int32_t)x;
}
static inline int32_t lrc_y(uint32_t i, float slope) { return f2i_sat((float)i * slope); }

/* Residuals: GC/LRC/LRC_Generator/LRC_Generator_kernel.cu:68-88 (the search loop there
 * terminates at docID - y); min shift: LRC_Generator.cu:165-180; blocks of `seg`
 * with ParaPFor: GC/LRC/LRC_Compression/main.cpp:163-199. */
uint64_t orc_lrc_encode_list(const uint32_t *L, uint32_t n, uint32_t seg, float frac,
                             orc_lrc_info *info, uint32_t *words,
                             uint32_t *seg_head, uint32_t *seg_off)
{
    orc_lrc_fit(L, n, info);
    if (n == 0) return 0;
    uint32_t *vd = (uint32_t *)malloc((size_t)n * 4);
    if (n > 1) {
        int32_t mn = 0;
        for (uint32_t i = 0; i < n; ++i) {
            int32_t d = (int32_t)(L[i] - ((uint32_t)lrc_y(i, info->fSlope) + (uint32_t)info->iIntercept));
            vd[i] = (uint32_t)d;
            if (i == 0 || d < mn) mn = d;
        }
        for (uint32_t i = 0; i < n; ++i) vd[i] = (uint32_t)((int32_t)vd[i] - mn);
        info->iIntercept += mn;
    } else {
        vd[0] = 0;
    }
    uint64_t w = 0;
    uint32_t nseg = (uint32_t)div_roundup(n, seg);
    for (uint32_t s = 0; s < nseg; ++s) {
        uint32_t st = s * seg, len = (s + 1 < nseg) ? seg : n - st;
        seg_head[s] = L[st];
        seg_off[s] = (uint32_t)w;
        uint32_t k = orc_pfor_encode(vd + st, len, frac, words + w);
        if (!k) { free(vd); return 0; }
        w += k;
    }
    free(vd);
    return w;
}

/* GC/LRC/LRC_Decompression/LRC_Decompression_kernel.cu:206-256:
 * docID[i] = VD[i] + int((float)i * slope) + icpt with i the list-global index. */
int orc_lrc_decode_list(const uint32_t *words, uint32_t n, uint32_t seg,
                        const orc_lrc_info *info, const uint32_t *seg_off, uint32_t *out)
{
    uint32_t nseg = (uint32_t)div_roundup(n, seg);
    for (uint32_t s = 0; s < nseg; ++s) {
        uint32_t st = s * seg, len = (s + 1 < nseg) ? seg : n - st;
        orc_pfor_decode(words + seg_off[s], len, out + st);
        for (uint32_t k = 0; k < len; ++k)
            out[st + k] += ((uint32_t)lrc_y(st + k, info->fSlope) + (uint32_t)info->iIntercept);
    }
    return 0;
}

/* --------------------------------------------------------------- SegLRC */

/* GC/SegLRC/SegLRC_Generator/SegLRC_Generator.cu:264-331 (same routine in
 * HS_SegLRC_Generator.cu:292-359): slope/intercept rounded to float, safe range
 * computed in float. */
void orc_seg_fit(const uint32_t *L, uint32_t n, orc_seg_info *info)
{
    if (n == 0) { info->fSlope = 1; info->fIntercept = 0; info->nRangeLeft = 0; info->nRangeRight = 0; return; }
    if (n == 1) { info->fSlope = 1; info->fIntercept = (float)L[0]; info->nRangeLeft = 0; info->nRangeRight = 0; return; }
    double xa = 0, ya = 0;
    for (uint32_t i = 0; i < n; ++i) { xa += i; ya += L[i]; }
    xa /= n; ya /= n;
    double sx = 0, sxy = 0;
    for (uint32_t i = 0; i < n; ++i) {
        double y = (double)L[i];
        sx += (double)((i - xa) * (i - xa));
        sxy += (double)((i - xa) * (y - ya));
    }
    info->fSlope = (float)(sxy / sx);
    info->fIntercept = (float)(ya - (double)info->fSlope * xa);
    float rl = 0, rr = 0;
    for (uint32_t i = 0; i < n; ++i) {
        float px = ((float)L[i] - info->fIntercept) / info->fSlope;
        float dv = px - (float)i;
        if (dv > rl) rl = dv;
        else if (dv < rr) rr = dv;
    }
    info->nRangeLeft = (uint32_t)rl + 1;
    info->nRangeRight = (uint32_t)(0 - rr) + 1;
}

static inline int32_t seg_y(uint32_t j, float s, float b) { return f2i_sat(fmaf((float)j, s, b)); }

/* residual against a segment fit, then per-unit min shift ("base"):
 * SegLRC_Generator_kernel.cu:62-100, SegLRC_Generator.cu:185-233. */
static int32_t seg_residuals(const uint32_t *L, uint32_t len, const orc_seg_info *f, uint32_t *vd)
{
    int32_t mn = 0;
    for (uint32_t j = 0; j < len; ++j) {
        int32_t d = (int32_t)(L[j] - (uint32_t)seg_y(j, f->fSlope, f->fIntercept));
        vd[j] = (uint32_t)d;
        if (j == 0 || d < mn) mn = d;
    }
    for (uint32_t j = 0; j < len; ++j) vd[j] = (uint32_t)((int32_t)vd[j] - mn);
    return len ? mn : 0;
}

uint64_t orc_seglrc_encode_list(const uint32_t *L, uint32_t n, uint32_t seg, float frac,
                                orc_seg_info *seg_fit, int32_t *base, uint32_t *words,
                                uint32_t *seg_head, uint32_t *seg_off, orc_seg_info *list_fit)
{
    if (list_fit) orc_seg_fit(L, n, list_fit);
    if (n == 0) return 0;
    uint32_t *vd = (uint32_t *)malloc((size_t)seg * 4);
    uint64_t w = 0;
    uint32_t nseg = (uint32_t)div_roundup(n, seg);
    for (uint32_t s = 0; s < nseg; ++s) {
        uint32_t st = s * seg, len = (s + 1 < nseg) ? seg : n - st;
        orc_seg_fit(L + st, len, &seg_fit[s]);
        base[s] = seg_residuals(L + st, len, &seg_fit[s], vd);
        seg_head[s] = L[st];
        seg_off[s] = (uint32_t)w;
        uint32_t k = orc_pfor_encode(vd, len, frac, words + w);
        if (!k) { free(vd); return 0; }
        w += k;
    }
    free(vd);
    return w;
}

/* GC/SegLRC/SegLRC_Decompression/SegLRC_Decompression_kernel.cu:200-256:
 * docID = VD + (int)(j * slope + icpt) + base, j in-segment. */
static void seg_block_decode(const uint32_t *blk, uint32_t len, const orc_seg_info *f, int32_t b, uint32_t *out)
{
    orc_pfor_decode(blk, len, out);
    for (uint32_t j = 0; j < len; ++j)
        out[j] += ((uint32_t)seg_y(j, f->fSlope, f->fIntercept) + (uint32_t)b);
}

int orc_seglrc_decode_list(const uint32_t *words, uint32_t n, uint32_t seg,
                           const orc_seg_info *seg_fit, const int32_t *base,
                           const uint32_t *seg_off, uint32_t *out)
{
    uint32_t nseg = (uint32_t)div_roundup(n, seg);
    for (uint32_t s = 0; s < nseg; ++s) {
        uint32_t st = s * seg, len = (s + 1 < nseg) ? seg : n - st;
        seg_block_decode(words + seg_off[s], len, &seg_fit[s], base[s], out + st);
    }
    return 0;
}

/* ------------------------------------------------------------------- HS */

/* GC/HS_SegLRC/HS_SegLRC_Generator/HS_SegLRC_Generator.cu:503-594. */
uint32_t orc_hs_table(const uint32_t *L, uint32_t n, uint32_t bucket_target,
                      uint32_t fixed_shift, uint32_t *info)
{
    uint32_t maxb = n ? gccbits(L[n - 1]) : 0;
    if (fixed_shift == 0 && n <= bucket_target) {
        info[0] = maxb; info[1] = 1; info[2] = 0; info[3] = n;
        return 4;
    }
    uint32_t shift;
    if (fixed_shift) {
        shift = fixed_shift;
    } else {
        uint32_t nb_exp = (uint32_t)div_roundup(n, bucket_target);
        shift = maxb - gccbits(nb_exp);
    }
    info[0] = shift;
    uint32_t cur = 0;
    info[2] = 0;
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t b = L[i] >> shift;
        while (cur < b) { ++cur; info[2 + cur] = i; }
    }
    ++cur;
    info[2 + cur] = n;
    info[1] = cur;
    return cur + 3;
}

/* Uncompressed hash-segment table, RunHashSegInfoGenerator of
 * src/gpu/Uncompressed/src/cuda_hashSegment_new/cuda_hashSegment.cu:815-947:
 * DOCID_BITNUM = log(max)/log(2) rounded up when it has a fraction (:872-881);
 * lists of at most UNI_LEN (= target) docIDs get [bits, 0, n] (:884-891);
 * longer ones nK = log(n / target)/log(2) rounded up the same way, shift =
 * bits - nK, then the walk that writes each bucket's first index and, after
 * the last bucket, n (:892-922): [shift][start_0 .. start_nb], nb = (max >>
 * shift) + 1.  A fixed shift (not in the reference: the C3 layout on the top
 * docID bits) walks with that shift for every list.  Where the reference takes
 * log(0) (max = 0, or an empty list, whose "last" docID it reads past the list)
 * this restatement uses bits = 0 and, for an empty list, [shift, 0, 0]. */
static uint32_t gu_ceil_log2(double x)
{
    if (!(x > 0.0)) return 0;
    double d = log(x) / log((double)2);
    if (d <= 0.0) return 0;
    if (d - (unsigned int)d > 0) return ((unsigned int)d) + 1;
    return (unsigned int)d;
}

uint32_t orc_gu_hashseg_table(const uint32_t *L, uint32_t n, uint32_t target, uint32_t fixed_shift,
                              uint32_t *info)
{
    if (n == 0) { info[0] = fixed_shift; info[1] = 0; info[2] = 0; return 3; }
    uint32_t shift = fixed_shift;
    if (!fixed_shift) {
        uint32_t bits = gu_ceil_log2((double)L[n - 1]);
        if (n <= target) { info[0] = bits; info[1] = 0; info[2] = n; return 3; }
        shift = bits - gu_ceil_log2(((double)n) / target);
    }
    info[0] = shift;
    uint32_t cur = 0;
    info[1] = 0;
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t b = L[i] >> shift;
        while (cur < b) { ++cur; info[1 + cur] = i; }
    }
    ++cur;
    info[1 + cur] = n;
    return cur + 2;
}

/* per-bucket fit, residual, base: HS_SegLRC_Generator.cu:95-288 and
 * HS_SegLRC_Generator_kernel.cu:64-114; packing: HS_SegLRC_Compression/main.cpp:149-178
 * (an empty bucket has no block; its segOffset equals the next block's). */
uint64_t orc_hs_encode_list(const uint32_t *L, uint32_t n, const uint32_t *info, float frac,
                            orc_seg_info *fit, int32_t *base, uint32_t *words, uint32_t *seg_off)
{
    uint32_t nb = info[1];
    const uint32_t *st = info + 2;
    uint32_t mx = 0;
    for (uint32_t b = 0; b < nb; ++b) if (st[b + 1] - st[b] > mx) mx = st[b + 1] - st[b];
    uint32_t *vd = (uint32_t *)malloc((size_t)(mx ? mx : 1) * 4);
    uint64_t w = 0;
    (void)n;
    for (uint32_t b = 0; b < nb; ++b) {
        uint32_t len = st[b + 1] - st[b];
        orc_seg_fit(L + st[b], len, &fit[b]);
        base[b] = seg_residuals(L + st[b], len, &fit[b], vd);
        seg_off[b] = (uint32_t)w;
        if (len > 0) {
            uint32_t k = orc_pfor_encode(vd, len, frac, words + w);
            if (!k) { free(vd); return 0; }
            w += k;
        }
    }
    free(vd);
    return w;
}

int orc_hs_decode_list(const uint32_t *words, uint32_t n, const uint32_t *info,
                       const orc_seg_info *fit, const int32_t *base,
                       const uint32_t *seg_off, uint32_t *out)
{
    uint32_t nb = info[1];
    const uint32_t *st = info + 2;
    (void)n;
    for (uint32_t b = 0; b < nb; ++b) {
        uint32_t len = st[b + 1] - st[b];
        if (len) seg_block_decode(words + seg_off[b], len, &fit[b], base[b], out + st[b]);
    }
    return 0;
}

/* -------------------------------------------------------------- ParaPFD */

/* GC/ParaPFD/ParaPFD_Compression/main.cpp:151-189 */
uint64_t orc_pfd_encode_list(const uint32_t *L, uint32_t n, uint32_t seg, float frac,
                             uint32_t *words, uint32_t *seg_head, uint32_t *seg_median,
                             uint32_t *seg_off)
{
    if (n == 0) return 0;
    uint32_t *g = (uint32_t *)malloc((size_t)n * 4);
    g[0] = L[0];
    for (uint32_t j = 1; j < n; ++j) g[j] = L[j] - L[j - 1];
    uint64_t w = 0;
    uint32_t nseg = (uint32_t)div_roundup(n, seg);
    for (uint32_t s = 0; s < nseg; ++s) {
        uint32_t st = s * seg, len = (s + 1 < nseg) ? seg : n - st;
        seg_head[s] = L[st];
        seg_median[s] = L[st + len / 2];
        seg_off[s] = (uint32_t)w;
        uint32_t k = orc_pfor_encode(g + st, len, frac, words + w);
        if (!k) { free(g); return 0; }
        w += k;
    }
    free(g);
    return w;
}

/* GC/ParaPFD/ParaPFD_Decompression/ParaPFD_Decompression_kernel.cu:198-271:
 * unpack, patch exceptions, replace gap 0 by segHead, inclusive prefix sum. */
static void pfd_block_decode(const uint32_t *blk, uint32_t len, uint32_t head, uint32_t *out)
{
    orc_pfor_decode(blk, len, out);
    out[0] = head;
    for (uint32_t j = 1; j < len; ++j) out[j] += out[j - 1];
}

int orc_pfd_decode_list(const uint32_t *words, uint32_t n, uint32_t seg,
                        const uint32_t *seg_head, const uint32_t *seg_off, uint32_t *out)
{
    uint32_t nseg = (uint32_t)div_roundup(n, seg);
    for (uint32_t s = 0; s < nseg; ++s) {
        uint32_t st = s * seg, len = (s + 1 < nseg) ? seg : n - st;
        pfd_block_decode(words + seg_off[s], len, seg_head[s], out + st);
    }
    return 0;
}

/* ------------------------------------------------------------- NewPFD */

/* Simple16 field widths per selector, slot by slot (GI/Simple16.h:16-34). */
static const uint8_t S16_W[16][28] = {
    {1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1},
    {2,2,2,2,2,2,2,1,1,1,1,1,1,1,1,1,1,1,1,1,1},
    {1,1,1,1,1,1,1,2,2,2,2,2,2,2,1,1,1,1,1,1,1},
    {1,1,1,1,1,1,1,1,1,1,1,1,1,1,2,2,2,2,2,2,2},
    {2,2,2,2,2,2,2,2,2,2,2,2,2,2},
    {4,3,3,3,3,3,3,3,3},
    {3,4,4,4,4,3,3,3},
    {4,4,4,4,4,4,4},
    {5,5,5,5,4,4},
    {4,4,5,5,5,5},
    {6,6,6,5,5},
    {5,5,6,6,6},
    {7,7,7,7},
    {10,9,9},
    {14,14},
    {28}};
static const uint8_t S16_N[16] = {28, 21, 21, 21, 14, 9, 8, 7, 6, 6, 5, 5, 4, 3, 2, 1};

/* GI/Simple16.h:13-48: per word, the first selector whose next min(count,
 * left) slots hold the values.  Returns words written, 0 if a value needs
 * more than 28 bits. */
uint32_t orc_s16_encode(const uint32_t *in, uint32_t n, uint32_t *out)
{
    uint32_t w = 0;
    for (uint32_t i = 0; i < n; ++w) {
        uint32_t k = 0;
        for (; k < 16; ++k) {
            uint32_t vol = S16_N[k] < n - i ? S16_N[k] : n - i, j = 0, sh = 0, word = k << 28;
            for (; j < vol && in[i + j] < (1u << S16_W[k][j]); ++j) {
                word += in[i + j] << sh;
                sh += S16_W[k][j];
            }
            if (j == vol) {
                out[w] = word;
                i += vol;
                break;
            }
        }
        if (k == 16) return 0;
    }
    return w;
}

/* S16_DECODE (GI/Simple16.h:52-258): whole words until at least n values;
 * writes up to n + 27 values, returns words consumed. */
uint32_t orc_s16_decode(const uint32_t *in, uint32_t n, uint32_t *out)
{
    uint32_t w = 0, got = 0;
    while (got < n) {
        uint32_t word = in[w++], k = word >> 28, sh = 0;
        for (uint32_t j = 0; j < S16_N[k]; ++j) {
            out[got++] = (word >> sh) & ((1u << S16_W[k][j]) - 1u);
            sh += S16_W[k][j];
        }
    }
    return w;
}

/* GI/NewPFor.cpp:54-99.  b as ParaPFor's rule, raised until every high part
 * fits 28 bits (where the reference's encoder would not return).  Returns
 * words written, 0 if unrepresentable. */
uint32_t orc_npfor_encode(const uint32_t *in, uint32_t n, float frac, uint32_t *out)
{
    uint32_t b = orc_pfor_best_b(in, n, frac), mb = maxbits(in, n);
    if (mb > 28 && b < mb - 28) b = mb - 28;
    if (b > 31) return 0;
    uint32_t *low = (uint32_t *)malloc(((size_t)n + 1) * 4);
    uint32_t *all = (uint32_t *)malloc((2 * (size_t)n + 1) * 4);
    uint32_t *pos = (uint32_t *)malloc(((size_t)n + 1) * 4);
    uint32_t en = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (gccbits(in[i]) > b) {
            low[i] = in[i] & ((1u << b) - 1u);
            all[n + en] = in[i] >> b;
            pos[en++] = i;
        } else {
            low[i] = in[i];
        }
    }
    out[0] = (b << 26) + en;
    pack_bits(low, n, b, out + 1);
    uint32_t w = 1 + (uint32_t)div_roundup((uint64_t)n * b, 32);
    if (en) {
        /* all = [pos0, pos1 - pos0 - 1, ..., highs] */
        for (uint32_t i = 0; i < en; ++i) all[i] = i ? pos[i] - pos[i - 1] - 1 : pos[0];
        memmove(all + en, all + n, (size_t)en * 4);
        uint32_t k = orc_s16_encode(all, 2 * en, out + w);
        w = k ? w + k : 0;
    }
    free(low);
    free(all);
    free(pos);
    return w;
}

/* GI/NewPFor.cpp:103-133.  Returns words consumed. */
uint32_t orc_npfor_decode(const uint32_t *in, uint32_t n, uint32_t *out)
{
    uint32_t d = in[0], b = (d >> 26) % 32, en = d & 65535u;
    unpack_bits(in + 1, n, b, out);
    uint32_t w = 1 + (uint32_t)div_roundup((uint64_t)n * b, 32);
    if (en) {
        uint32_t *all = (uint32_t *)malloc((2 * (size_t)en + 28) * 4);
        w += orc_s16_decode(in + w, 2 * en, all);
        uint32_t p = all[0];
        for (uint32_t i = 0; i < en; ++i) {
            if (p < n) out[p] += all[en + i] << b;
            if (i + 1 < en) p += all[i + 1] + 1;
        }
        free(all);
    }
    return w;
}

/* GC/NewPFD/NewPFD_Compression/main.cpp:151-189: d-gaps chained across
 * blocks (gap 0 = docID 0), 256- (or seg-) element NewPFor blocks. */
uint64_t orc_npfd_encode_list(const uint32_t *L, uint32_t n, uint32_t seg, float frac,
                              uint32_t *words, uint32_t *seg_head, uint32_t *seg_median,
                              uint32_t *seg_off)
{
    if (n == 0) return 0;
    uint32_t *g = (uint32_t *)malloc((size_t)n * 4);
    g[0] = L[0];
    for (uint32_t j = 1; j < n; ++j) g[j] = L[j] - L[j - 1];
    uint64_t w = 0;
    uint32_t nseg = (uint32_t)div_roundup(n, seg);
    for (uint32_t s = 0; s < nseg; ++s) {
        uint32_t st = s * seg, len = (s + 1 < nseg) ? seg : n - st;
        seg_head[s] = L[st];
        seg_median[s] = L[st + len / 2];
        seg_off[s] = (uint32_t)w;
        uint32_t k = orc_npfor_encode(g + st, len, frac, words + w);
        if (!k) { free(g); return 0; }
        w += k;
    }
    free(g);
    return w;
}

/* GC/NewPFD/NewPFD_Decompression/NewPFD_Decompression_kernel.cu:405-470:
 * unpack + patch, gap 0 replaced by segHead, inclusive prefix sum. */
int orc_npfd_decode_list(const uint32_t *words, uint32_t n, uint32_t seg,
                         const uint32_t *seg_head, const uint32_t *seg_off, uint32_t *out)
{
    uint32_t nseg = (uint32_t)div_roundup(n, seg);
    for (uint32_t s = 0; s < nseg; ++s) {
        uint32_t st = s * seg, len = (s + 1 < nseg) ? seg : n - st;
        orc_npfor_decode(words + seg_off[s], len, out + st);
        out[st] = seg_head[s];
        for (uint32_t j = 1; j < len; ++j) out[st + j] += out[st + j - 1];
    }
    return 0;
}

/* ------------------------------------------------------- index + queries */

static uint32_t term_nseg(const orc_index *ix, uint32_t t)
{
    if (ix->codec == ORC_HS) return ix->hash_info[ix->hash_off[t] + 1];
    return (uint32_t)div_roundup(ix->ucnt[t], ix->seg);
}

static uint64_t term_words(const orc_index *ix, uint32_t t)
{
    uint64_t end = (t + 1 < ix->nterms) ? ix->uwlist[t + 1] : ix->total_words;
    return end - ix->uwlist[t];
}

int orc_decode_term(const orc_index *ix, uint32_t t, uint32_t *out)
{
    uint32_t n = ix->ucnt[t];
    const uint32_t *w = ix->words + ix->uwlist[t];
    uint32_t sn = ix->segnum[t];
    switch (ix->codec) {
    case ORC_LRC: return orc_lrc_decode_list(w, n, ix->seg, &ix->lrc[t], ix->seg_off + sn, out);
    case ORC_SEGLRC: return orc_seglrc_decode_list(w, n, ix->seg, ix->fit + sn, ix->base + sn, ix->seg_off + sn, out);
    case ORC_HS: return orc_hs_decode_list(w, n, ix->hash_info + ix->hash_off[t], ix->fit + sn, ix->base + sn, ix->seg_off + sn, out);
    case ORC_PFD: return orc_pfd_decode_list(w, n, ix->seg, ix->seg_head + sn, ix->seg_off + sn, out);
    case ORC_NPFD: return orc_npfd_decode_list(w, n, ix->seg, ix->seg_head + sn, ix->seg_off + sn, out);
    }
    return -1;
}

/* Decode block s of term t into out; returns its length. */
static uint32_t decode_block(const orc_index *ix, uint32_t t, uint32_t s, uint32_t *out)
{
    uint32_t n = ix->ucnt[t], sn = ix->segnum[t];
    const uint32_t *w = ix->words + ix->uwlist[t];
    if (ix->codec == ORC_HS) {
        const uint32_t *st = ix->hash_info + ix->hash_off[t] + 2;
        uint32_t len = st[s + 1] - st[s];
        if (len) seg_block_decode(w + ix->seg_off[sn + s], len, &ix->fit[sn + s], ix->base[sn + s], out);
        return len;
    }
    uint32_t nseg = (uint32_t)div_roundup(n, ix->seg);
    uint32_t st = s * ix->seg, len = (s + 1 < nseg) ? ix->seg : n - st;
    const uint32_t *blk = w + ix->seg_off[sn + s];
    switch (ix->codec) {
    case ORC_LRC:
        orc_pfor_decode(blk, len, out);
        for (uint32_t k = 0; k < len; ++k)
            out[k] += ((uint32_t)lrc_y(st + k, ix->lrc[t].fSlope) + (uint32_t)ix->lrc[t].iIntercept);
        break;
    case ORC_SEGLRC: seg_block_decode(blk, len, &ix->fit[sn + s], ix->base[sn + s], out); break;
    case ORC_PFD: pfd_block_decode(blk, len, ix->seg_head[sn + s], out); break;
    }
    return len;
}

/* Block of term t whose docID range may hold v; -1 if none can. */
static int64_t find_block(const orc_index *ix, uint32_t t, uint32_t v)
{
    if (ix->ucnt[t] == 0) return -1;
    if (ix->codec == ORC_HS) {
        const uint32_t *h = ix->hash_info + ix->hash_off[t];
        uint32_t b = v >> h[0];
        return (b < h[1]) ? (int64_t)b : -1;
    }
    const uint32_t *head = ix->seg_head + ix->segnum[t];
    uint32_t nseg = term_nseg(ix, t);
    /* last segment with head <= v */
    uint32_t lo = 0, hi = nseg;
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (head[mid] <= v) lo = mid + 1; else hi = mid;
    }
    return (int64_t)lo - 1;
}

static int bsearch_u32(const uint32_t *a, uint32_t n, uint32_t v)
{
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (a[mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo < n && a[lo] == v;
}

uint64_t orc_intersect2(const uint32_t *a, uint64_t na, const uint32_t *b, uint64_t nb, uint32_t *out)
{
    uint64_t i = 0, j = 0, k = 0;
    while (i < na && j < nb) {
        if (a[i] < b[j]) ++i;
        else if (b[j] < a[i]) ++j;
        else { out[k++] = a[i]; ++i; ++j; }
    }
    return k;
}

/* Terms of a query ordered by ascending length with a stable insertion sort,
 * as CPUPreprocess does (LRC_Intersection.cu:157-179). */
static void sort_terms(const orc_index *ix, const uint32_t *in, uint32_t tc, uint32_t *out)
{
    for (uint32_t i = 0; i < tc; ++i) {
        uint32_t t = in[i], k = i;
        while (k && ix->ucnt[t] < ix->ucnt[out[k - 1]]) { out[k] = out[k - 1]; --k; }
        out[k] = t;
    }
}

static uint32_t max_block_len(const orc_index *ix, uint32_t t)
{
    if (ix->codec != ORC_HS) return ix->seg;
    const uint32_t *h = ix->hash_info + ix->hash_off[t];
    uint32_t mx = 0;
    for (uint32_t b = 0; b < h[1]; ++b) if (h[3 + b] - h[2 + b] > mx) mx = h[3 + b] - h[2 + b];
    return mx;
}

/* One query; returns result count; results in *buf (ascending). */
static uint32_t eval_one(const orc_index *ix, const uint32_t *terms, uint32_t tc, uint32_t **buf, uint32_t *bufcap)
{
    if (tc == 0) return 0;
    uint32_t st[256];
    uint32_t *srt = tc <= 256 ? st : (uint32_t *)malloc((size_t)tc * 4);
    sort_terms(ix, terms, tc, srt);
    uint32_t n0 = ix->ucnt[srt[0]];
    if (*bufcap < n0 + 1) { free(*buf); *bufcap = n0 + 1; *buf = (uint32_t *)malloc((size_t)*bufcap * 4); }
    uint32_t *a = *buf;
    orc_decode_term(ix, srt[0], a);
    uint32_t na = n0;
    for (uint32_t j = 1; j < tc && na; ++j) {
        uint32_t t = srt[j];
        uint32_t mbl = max_block_len(ix, t);
        uint32_t *blk = (uint32_t *)malloc((size_t)(mbl ? mbl : 1) * 4);
        int64_t cached = -1;
        uint32_t blen = 0, out = 0;
        for (uint32_t i = 0; i < na; ++i) {
            uint32_t v = a[i];
            int64_t b = find_block(ix, t, v);
            if (b < 0) continue;
            if (b != cached) { blen = decode_block(ix, t, (uint32_t)b, blk); cached = b; }
            if (bsearch_u32(blk, blen, v)) a[out++] = v;
        }
        free(blk);
        na = out;
    }
    if (srt != st) free(srt);
    return na;
}

typedef struct {
    const orc_index *ix;
    const uint32_t *qterms;
    const uint64_t *qoff;
    uint32_t q0, q1;
    uint32_t *counts;
    uint32_t *res;
    uint64_t nres, cap;
    double *q_ms;
    int err;
} eval_job;

static double now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

static void *eval_worker(void *arg)
{
    eval_job *jb = (eval_job *)arg;
    uint32_t *buf = NULL, cap = 0;
    jb->cap = 1 << 16;
    jb->res = (uint32_t *)malloc(jb->cap * 4);
    jb->nres = 0;
    for (uint32_t q = jb->q0; q < jb->q1; ++q) {
        double t0 = jb->q_ms ? now_ms() : 0;
        uint64_t o = jb->qoff[q];
        uint32_t tc = (uint32_t)(jb->qoff[q + 1] - o);
        uint32_t c = eval_one(jb->ix, jb->qterms + o, tc, &buf, &cap);
        if (jb->nres + c > jb->cap) {
            while (jb->nres + c > jb->cap) jb->cap *= 2;
            jb->res = (uint32_t *)realloc(jb->res, jb->cap * 4);
        }
        memcpy(jb->res + jb->nres, buf, (size_t)c * 4);
        jb->nres += c;
        jb->counts[q] = c;
        if (jb->q_ms) jb->q_ms[q] = now_ms() - t0;
    }
    free(buf);
    return NULL;
}

int orc_eval_queries(const orc_index *ix, const uint32_t *qterms, const uint64_t *qoff,
                     uint32_t nq, int nthreads, uint32_t *cum_counts, uint32_t *results,
                     uint64_t res_cap, double *q_ms)
{
    if (nthreads < 1) nthreads = 1;
    if ((uint32_t)nthreads > nq) nthreads = nq ? (int)nq : 1;
    eval_job *jobs = (eval_job *)calloc((size_t)nthreads, sizeof(eval_job));
    pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
    uint32_t *counts = (uint32_t *)calloc(nq ? nq : 1, 4);
    for (int t = 0; t < nthreads; ++t) {
        jobs[t].ix = ix; jobs[t].qterms = qterms; jobs[t].qoff = qoff;
        jobs[t].q0 = (uint32_t)((uint64_t)nq * t / nthreads);
        jobs[t].q1 = (uint32_t)((uint64_t)nq * (t + 1) / nthreads);
        jobs[t].counts = counts; jobs[t].q_ms = q_ms;
        if (nthreads > 1) pthread_create(&th[t], NULL, eval_worker, &jobs[t]);
        else eval_worker(&jobs[t]);
    }
    if (nthreads > 1) for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    int rc = 0;
    uint64_t pos = 0;
    for (int t = 0; t < nthreads; ++t) {
        if (pos + jobs[t].nres > res_cap) rc = -1;
        else if (results) memcpy(results + pos, jobs[t].res, jobs[t].nres * 4);
        pos += jobs[t].nres;
        free(jobs[t].res);
    }
    uint64_t acc = 0;
    for (uint32_t q = 0; q < nq; ++q) { acc += counts[q]; cum_counts[q] = (uint32_t)acc; }
    free(counts); free(jobs); free(th);
    return rc;
}

int orc_eval_queries_full(const orc_index *ix, const uint32_t *qterms, const uint64_t *qoff,
                          uint32_t nq, uint32_t *cum_counts, uint32_t *results, uint64_t res_cap)
{
    uint64_t pos = 0;
    for (uint32_t q = 0; q < nq; ++q) {
        uint64_t o = qoff[q];
        uint32_t tc = (uint32_t)(qoff[q + 1] - o);
        uint64_t na = 0;
        uint32_t *a = NULL;
        for (uint32_t j = 0; j < tc; ++j) {
            uint32_t t = qterms[o + j], n = ix->ucnt[t];
            uint32_t *b = (uint32_t *)malloc((size_t)(n ? n : 1) * 4);
            orc_decode_term(ix, t, b);
            if (j == 0) { a = b; na = n; continue; }
            na = orc_intersect2(a, na, b, n, a);
            free(b);
        }
        if (pos + na > res_cap) { free(a); return -1; }
        if (na) memcpy(results + pos, a, na * 4);
        free(a);
        pos += na;
        cum_counts[q] = (uint32_t)pos;
    }
    return 0;
}

/* LRC_Intersection.cu:354-379 */
void orc_checksum(const uint32_t *cum, const uint32_t *r, uint32_t nq, uint64_t cs[3])
{
    cs[0] = cs[1] = cs[2] = 0;
    for (uint32_t i = 0; i < nq; ++i) {
        uint32_t b = i ? cum[i - 1] : 0, c = cum[i] - b;
        if (c) {
            cs[1] += c;
            cs[2] ^= r[b];
            cs[2] ^= r[b + c - 1];
            cs[2] ^= r[b + c / 2];
        } else {
            ++cs[0];
        }
    }
}

static uint32_t meta_bytes(const orc_index *ix)
{
    switch (ix->codec) {
    case ORC_LRC: return 8;     /* segOffset + segHead */
    case ORC_SEGLRC: return 28; /* segOffset + segHead + fit(16) + base */
    case ORC_HS: return 32;     /* segOffset + fit(16) + base + 2 bucket starts */
    case ORC_PFD: return 8;     /* segOffset + segHead */
    }
    return 0;
}

static uint64_t block_words(const orc_index *ix, uint32_t t, uint32_t s)
{
    uint32_t sn = ix->segnum[t], ns = term_nseg(ix, t);
    uint64_t end = (s + 1 < ns) ? ix->seg_off[sn + s + 1] : term_words(ix, t);
    return end - ix->seg_off[sn + s];
}

uint64_t orc_query_bytes(const orc_index *ix, const uint32_t *qterms, const uint64_t *qoff,
                         uint32_t nq, uint64_t *full_bytes)
{
    uint64_t tot = 0, full = 0;
    uint32_t *buf = NULL, cap = 0;
    uint32_t mb = meta_bytes(ix);
    for (uint32_t q = 0; q < nq; ++q) {
        uint64_t o = qoff[q];
        uint32_t tc = (uint32_t)(qoff[q + 1] - o);
        if (!tc) continue;
        uint32_t srt[256];
        if (tc > 256) return 0;
        sort_terms(ix, qterms + o, tc, srt);
        for (uint32_t j = 0; j < tc; ++j)
            full += term_words(ix, srt[j]) * 4 + (uint64_t)term_nseg(ix, srt[j]) * mb + (ix->codec == ORC_LRC ? 12 : 0);
        uint32_t n0 = ix->ucnt[srt[0]];
        tot += term_words(ix, srt[0]) * 4 + (uint64_t)term_nseg(ix, srt[0]) * mb + (ix->codec == ORC_LRC ? 12 : 0);
        if (cap < n0 + 1) { free(buf); cap = n0 + 1; buf = (uint32_t *)malloc((size_t)cap * 4); }
        orc_decode_term(ix, srt[0], buf);
        uint32_t na = n0;
        for (uint32_t j = 1; j < tc && na; ++j) {
            uint32_t t = srt[j];
            uint32_t mbl = max_block_len(ix, t);
            uint32_t *blk = (uint32_t *)malloc((size_t)(mbl ? mbl : 1) * 4);
            int64_t cached = -1;
            uint32_t blen = 0, out = 0;
            if (ix->codec == ORC_LRC) tot += 12;
            for (uint32_t i = 0; i < na; ++i) {
                uint32_t v = buf[i];
                int64_t b = find_block(ix, t, v);
                if (b < 0) continue;
                if (b != cached) {
                    blen = decode_block(ix, t, (uint32_t)b, blk);
                    cached = b;
                    tot += block_words(ix, t, (uint32_t)b) * 4 + mb;
                }
                if (bsearch_u32(blk, blen, v)) buf[out++] = v;
            }
            free(blk);
            na = out;
        }
        tot += (uint64_t)na * 4;
        full += (uint64_t)na * 4;
    }
    free(buf);
    if (full_bytes) *full_bytes = full;
    return tot;
}

/* ----------------------------------------------------------- learned models */

/* PLA-256: knots at t = p of t(v) = ((float)v - x0) * xs, xs = 256 / span;
 * y[p] = number of elements with t < p (the empirical CDF at the knots). */
void orc_pla_fit(const uint32_t *L, uint32_t n, float *x0, float *xs, float *y)
{
    *x0 = (float)L[0];
    *xs = 256.0f / ((float)L[n - 1] - *x0 + 1.0f);
    for (int p = 0; p < 257; ++p) {
        uint32_t c = 0;
        while (c < n && ((float)L[c] - *x0) * *xs < (float)p) ++c;
        y[p] = (float)c;
    }
}

float orc_pla_pred(const float *y, float x0, float xs, uint32_t v)
{
    float t = ((float)v - x0) * xs;
    t = fminf(fmaxf(t, 0.0f), 256.0f);
    int p = (int)t;
    if (p > 255) p = 255;
    const float f = t - (float)p;
    return fmaf(f, y[p + 1] - y[p], y[p]);
}

float orc_mlp_pred(const float *w, float x0, float xs, uint32_t n, uint32_t v)
{
    const float x = ((float)v - x0) * xs;
    float acc = w[96];
    for (int j = 0; j < 32; ++j) {
        const float h = fmaxf(fmaf(w[j], x, w[32 + j]), 0.0f);
        acc = fmaf(w[64 + j], h, acc);
    }
    return acc * (float)n;
}

/* PLA-Q: knot j at position q_j = floor(j (n - 1) / 256): x[j] = (float)L[q_j],
 * y[j] = q_j, s[j] = (y[j+1] - y[j]) / (x[j+1] - x[j]) (0 for coinciding
 * knots); params = x[257] y[257] s[256]. */
void orc_plaq_fit(const uint32_t *L, uint32_t n, float *prm)
{
    for (uint32_t j = 0; j < 257; ++j) {
        const uint32_t q = (uint32_t)(((uint64_t)j * (n - 1)) / 256u);
        prm[j] = (float)L[q];
        prm[257 + j] = (float)q;
    }
    for (uint32_t j = 0; j < 256; ++j) {
        const float dx = prm[j + 1] - prm[j];
        prm[514 + j] = dx > 0.0f ? (prm[257 + j + 1] - prm[257 + j]) / dx : 0.0f;
    }
}

/* the piece: linear scan for the last knot <= v (the device halves) */
float orc_plaq_pred(const float *prm, uint32_t v)
{
    const float fv = (float)v;
    if (fv < prm[0]) return prm[257];
    if (fv >= prm[256]) return prm[513];
    int p = 0;
    while (p + 1 < 256 && prm[p + 1] <= fv) ++p;
    return fmaf(fv - prm[p], prm[514 + p], prm[257 + p]);
}

float orc_model_max_err(int kind, const float *params, float x0, float xs, const uint32_t *L, uint32_t n)
{
    float e = 0.0f;
    for (uint32_t i = 0; i < n; ++i) {
        const float p = kind == 1   ? orc_pla_pred(params, x0, xs, L[i])
                        : kind == 3 ? orc_plaq_pred(params, L[i])
                                    : orc_mlp_pred(params, x0, xs, n, L[i]);
        e = fmaxf(e, fabsf(p - (float)i));
    }
    return e;
}
