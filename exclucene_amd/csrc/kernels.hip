// kernels.hip -- gfx950 kernels for compressed posting-list decode + exact
// k-term intersection (replaces the reference's *_Intersection, *_Decompression,
// saCompact and ucntResult kernels and the cudppScan call;
// GC = src/gpu/Compressed/src):
//   GC/LRC/LRC_Intersection/LRC_Intersection_kernel.cu:155-365,613-721
//   GC/SegLRC/SegLRC_Intersection/SegLRC_Intersection_kernel.cu:163-402
//   GC/HS_SegLRC/HS_SegLRC_Intersection/HS_SegLRC_Intersection_kernel.cu:162-404
//   GC/ParaPFD/ParaPFD_Intersection_FRAC/ParaPFD_Intersection_kernel.cu:145-458
//   GC/ParaPFD/ParaPFD_Decompression/ParaPFD_Decompression_kernel.cu:117-272
//
// Execution model (CDNA4, wave64):
//  * one wave owns one work unit = (query, one segment of its shortest list);
//    a 256-thread workgroup hosts 4 independent waves, each with a private LDS
//    slice, synchronised with wave barriers only.
//  * the unit's docIDs are decoded cooperatively into LDS and stay there as the
//    candidate set; each longer list (ascending length) filters it.
//  * per list: the wave brackets the segment range with a 64-ary search over
//    segHead (one gather + ballot per level), each candidate finds its segment
//    with 6 register shuffles (range <= 64 segments) or a binary search, then
//    either (tpe = 1) searches its segment by random-access bit extraction, or
//    (tpe = 64) the wave decodes each distinct needed segment into LDS with
//    coalesced loads and candidates binary-search LDS.
//  * survivors are compacted with ballot + mbcnt (order preserving), so the
//    next list only probes live candidates and no 0/1 flag array exists.
//  * results: each unit writes its survivors into the stage slot of its
//    shortest-list segment; a block scan over unit counts + a scatter produce
//    the contiguous per-query output and cumulative counts.
#include "exc_format.h"
#include "exc_internal.h"

namespace exc {
namespace {

constexpr int kWaves = 4;
constexpr uint32_t kNone = 0xffffffffu;

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint32_t uni(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint32_t below_me(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
__device__ __forceinline__ uint32_t popc(uint64_t m) { return (uint32_t)__popcll(m); }

// lb-bit value j of a packed run starting at p (LSB-first, may straddle two
// words; GI/unpack.h:14-23).  Reads p[w] and p[w+1]; every device array that
// is read this way carries kPadWords zero words at its end.
__device__ __forceinline__ uint32_t bits_at(const uint32_t* __restrict__ p, uint32_t lb, uint32_t j)
{
    const uint32_t bit = j * lb;
    const uint32_t* q = p + (bit >> 5);
    const uint64_t cw = (uint64_t)q[0] | ((uint64_t)q[1] << 32);
    return (uint32_t)(cw >> (bit & 31)) & ((1u << lb) - 1u);
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    const uint32_t ln = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o);
        if (ln >= (uint32_t)o) v += t;
    }
    return v;
}

// Number of elements <= v in the sorted array a[0..n) (v wave-uniform):
// 64-ary search, one gather + ballot per level.
__device__ uint32_t wave_upper_bound(const uint32_t* __restrict__ a, uint32_t n, uint32_t v)
{
    const uint32_t ln = lane_id();
    uint32_t lo = 0, hi = n;
    while (hi - lo > 64) {
        const uint32_t stride = (hi - lo + 63) >> 6;
        const uint32_t idx = lo + ln * stride;
        const bool p = idx < hi && a[idx] <= v;
        const uint32_t c = popc(__ballot(p));
        if (c == 0) return lo;
        const uint32_t nlo = lo + (c - 1) * stride + 1;
        const uint32_t nhi = min(lo + c * stride, hi);
        lo = uni(nlo);
        hi = uni(nhi);
    }
    const uint32_t idx = lo + ln;
    const bool p = idx < hi && a[idx] <= v;
    return lo + popc(__ballot(p));
}

// Same with a predicted window [wa, wb] for the answer; the window is used only
// if head[wa-1] <= v < head[wb] confirms it (exact fallback otherwise).
__device__ uint32_t wave_upper_bound_win(const uint32_t* __restrict__ a, uint32_t n, uint32_t v,
                                         uint32_t wa, uint32_t wb)
{
    const uint32_t ln = lane_id();
    wa = min(wa, n);
    wb = min(max(wb, wa), n);
    uint32_t x = 0;
    if (ln == 0 && wa > 0) x = a[wa - 1];
    if (ln == 1 && wb < n) x = a[wb];
    const uint32_t left = uni(__shfl(x, 0)), right = uni(__shfl(x, 1));
    const bool ok = (wa == 0 || left <= v) && (wb == n || right > v);
    if (!ok) return wave_upper_bound(a, n, v);
    return wa + wave_upper_bound(a + wa, wb - wa, v);
}

__device__ __forceinline__ uint32_t seg_len(const TermRec& R, uint32_t s, uint32_t seg)
{
    return (s + 1 == R.nseg) ? R.n - s * seg : seg;
}

// Decode segment s of list R into dst[0..len) with the whole wave (segment
// codecs).  Returns len.  Caller must wave_sync() before reading other lanes'.
template <int CODEC, int SEG>
__device__ uint32_t decode_seg_block(const DevIndex& ix, const TermRec& R, uint32_t s, uint32_t* dst)
{
    const uint32_t ln = lane_id();
    const uint32_t len = uni(seg_len(R, s, SEG));
    const uint32_t* blk = ix.words + R.woff + ix.seg_off[R.segnum + s];
    const uint32_t d = uni(blk[0]);
    const uint32_t lb = desc_lb(d);
    if constexpr (CODEC == EXC_LRC) {
        // GC/LRC/LRC_Decompression/LRC_Decompression_kernel.cu:206-256
#pragma unroll
        for (int i = 0; i < SEG / 64; ++i) {
            const uint32_t j = ln + 64 * i;
            if (j < len) dst[j] = bits_at(blk + 1, lb, j) + (uint32_t)(lrc_y(s * SEG + j, R.slope) + R.icpt);
        }
    } else if constexpr (CODEC == EXC_SEGLRC) {
        // GC/SegLRC/SegLRC_Decompression/SegLRC_Decompression_kernel.cu:200-256
        const float4 f = ix.fit[R.segnum + s];
        const int32_t b = ix.base[R.segnum + s];
#pragma unroll
        for (int i = 0; i < SEG / 64; ++i) {
            const uint32_t j = ln + 64 * i;
            if (j < len) dst[j] = bits_at(blk + 1, lb, j) + (uint32_t)(seg_y(j, f.x, f.y) + b);
        }
    } else {
        // ParaPFD: unpack, patch exceptions, head, inclusive scan
        // (GC/ParaPFD/ParaPFD_Decompression/ParaPFD_Decompression_kernel.cu:198-264).
        // Lane holds V consecutive gaps, so the scan is V-1 local adds plus one
        // wave scan of lane totals.
        constexpr int V = SEG / 64;
        uint32_t val[V];
#pragma unroll
        for (int t = 0; t < V; ++t) {
            const uint32_t j = ln * V + t;
            val[t] = (j < len) ? bits_at(blk + 1, lb, j) : 0u;
        }
        const uint32_t en = desc_en(d);
        if (en) {
#pragma unroll
            for (int t = 0; t < V; ++t) {
                const uint32_t j = ln * V + t;
                if (j < len) dst[j] = val[t];
            }
            wave_sync();
            const uint32_t eb = desc_eb(d), hb = desc_hb(d);
            const uint32_t* pp = blk + 1 + words_for(len, lb);
            const uint32_t* hp = pp + words_for(en, eb);
            for (uint32_t e = ln; e < en; e += 64) {
                const uint32_t p = bits_at(pp, eb, e);
                const uint32_t h = bits_at(hp, hb, e);
                if (p < len) dst[p] |= h << lb;
            }
            wave_sync();
#pragma unroll
            for (int t = 0; t < V; ++t) {
                const uint32_t j = ln * V + t;
                if (j < len) val[t] = dst[j];
            }
        }
        if (ln == 0) val[0] = ix.seg_head[R.segnum + s];
        uint32_t acc = 0;
#pragma unroll
        for (int t = 0; t < V; ++t) { acc += val[t]; val[t] = acc; }
        const uint32_t excl = wave_incl_scan(acc) - acc;
        wave_sync();
#pragma unroll
        for (int t = 0; t < V; ++t) {
            const uint32_t j = ln * V + t;
            if (j < len) dst[j] = val[t] + excl;
        }
    }
    return len;
}

// exact search of v in a random-access SegLRC / HS block of len values, with
// the per-block fit window [pivot - L, pivot + R] tried first and a full binary
// search when v may lie outside it.  The reference uses the window alone
// (HS_SegLRC_Intersection_kernel.cu:361-369, SegLRC_Intersection_kernel.cu:366-371),
// which drops results; the fallback keeps the answer exact.
__device__ __forceinline__ bool fit_block_search(const uint32_t* __restrict__ blk, uint32_t len, float4 f,
                                                 int32_t base, uint32_t v, bool use_window)
{
    const uint32_t lb = desc_lb(blk[0]);
    const uint32_t* p = blk + 1;
    int lo = 0, hi = (int)len - 1;
    if (use_window) {
        float fp = ((float)v - f.y) / f.x;
        fp = fminf(fmaxf(fp, -1.0e9f), 1.0e9f);
        const int64_t piv = (int64_t)fp;
        const int wl = (int)max(piv - (int64_t)__float_as_uint(f.z), (int64_t)0);
        const int wh = (int)min(piv + (int64_t)__float_as_uint(f.w), (int64_t)len - 1);
        if (wl <= wh) {
            int a = wl, c = wh;
            while (a <= c) {
                const int m = (a + c) >> 1;
                const uint32_t q = bits_at(p, lb, (uint32_t)m) + (uint32_t)(seg_y((uint32_t)m, f.x, f.y) + base);
                if (q == v) return true;
                if (q < v) a = m + 1; else c = m - 1;
            }
            if (a == wl && wl > 0) { lo = 0; hi = wl - 1; }
            else if (a == wh + 1 && wh < (int)len - 1) { lo = wh + 1; hi = (int)len - 1; }
            else return false;
        }
    }
    while (lo <= hi) {
        const int m = (lo + hi) >> 1;
        const uint32_t q = bits_at(p, lb, (uint32_t)m) + (uint32_t)(seg_y((uint32_t)m, f.x, f.y) + base);
        if (q == v) return true;
        if (q < v) lo = m + 1; else hi = m - 1;
    }
    return false;
}

// random-access search of v in segment s (LRC / SegLRC, no exceptions)
template <int CODEC, int SEG>
__device__ __forceinline__ bool ra_search(const ProbeArgs& A, const TermRec& R, uint32_t s, uint32_t v)
{
    const uint32_t len = seg_len(R, s, SEG);
    const uint32_t* blk = A.ix.words + R.woff + A.ix.seg_off[R.segnum + s];
    if constexpr (CODEC == EXC_SEGLRC) {
        return fit_block_search(blk, len, A.ix.fit[R.segnum + s], A.ix.base[R.segnum + s], v,
                                A.search != SEARCH_BINARY);
    } else {
        // GC/LRC/LRC_Intersection/LRC_Intersection_kernel.cu:333-355
        const uint32_t lb = desc_lb(blk[0]);
        int lo = 0, hi = (int)len - 1;
        while (lo <= hi) {
            const int m = (lo + hi) >> 1;
            const uint32_t q = bits_at(blk + 1, lb, (uint32_t)m) + (uint32_t)(lrc_y(s * SEG + (uint32_t)m, R.slope) + R.icpt);
            if (q == v) return true;
            if (q < v) lo = m + 1; else hi = m - 1;
        }
        return false;
    }
}

// Predicted window of "number of segment heads <= v" from the list-level fit.
template <int CODEC, int SEG>
__device__ __forceinline__ void fit_window(const TermRec& R, uint32_t v, uint32_t& wa, uint32_t& wb)
{
    float icpt = (CODEC == EXC_LRC) ? (float)R.icpt : __int_as_float(R.icpt);
    float p = ((float)v - icpt) / R.slope;
    p = fminf(fmaxf(p, -1.0e9f), 4.0e9f);
    const float plo = p - R.range_l - 1.0f, phi = p + R.range_r + 1.0f;
    wa = plo <= 0.0f ? 0u : (uint32_t)(plo / SEG);
    wb = phi <= 0.0f ? 1u : (uint32_t)(phi / SEG) + 2u;
}

// Filter cand[0..nlive) against list R (segment codecs).  Returns survivors.
template <int CODEC, int SEG>
__device__ uint32_t probe_seg(const ProbeArgs& A, const TermRec& R, uint32_t* cand, uint32_t nlive,
                              uint32_t* sarr, uint32_t* dlist, uint32_t* buf)
{
    constexpr int V = SEG / 64;
    constexpr int G = (512 / SEG) > 0 ? (512 / SEG) : 1;
    const uint32_t ln = lane_id();
    const uint32_t* head = A.ix.seg_head + R.segnum;
    const uint32_t vmin = uni(cand[0]), vmax = uni(cand[nlive - 1]);

    // ---- segment range of the candidate set: s_lo .. ub_hi-1
    uint32_t ub_lo, ub_hi;
    if (A.search != SEARCH_BINARY && R.nseg > 64) {
        uint32_t wa, wb;
        fit_window<CODEC, SEG>(R, vmin, wa, wb);
        ub_lo = wave_upper_bound_win(head, R.nseg, vmin, wa, wb);
        fit_window<CODEC, SEG>(R, vmax, wa, wb);
        ub_hi = wave_upper_bound_win(head, R.nseg, vmax, wa, wb);
    } else {
        ub_lo = wave_upper_bound(head, R.nseg, vmin);
        ub_hi = wave_upper_bound(head, R.nseg, vmax);
    }
    ub_lo = uni(ub_lo);
    ub_hi = uni(ub_hi);
    if (ub_hi == 0) return 0;  // every candidate precedes the list
    const uint32_t s_lo = ub_lo ? ub_lo - 1 : 0;
    const uint32_t rng = ub_hi - s_lo;
    const bool small = rng <= 64;

    // ---- segment of every candidate
    uint32_t hreg = kNone;
    if (small && ln < rng) hreg = head[s_lo + ln];
    const uint32_t h63 = uni(__builtin_amdgcn_readlane(hreg, 63));
    uint32_t v[V], s[V];
    bool live[V];
#pragma unroll
    for (int i = 0; i < V; ++i) {
        const uint32_t e = ln + 64 * i;
        live[i] = e < nlive;
        v[i] = live[i] ? cand[e] : kNone;
        uint32_t pos = 0;
        if (small) {
#pragma unroll
            for (uint32_t st = 32; st >= 1; st >>= 1) {
                const uint32_t h = __shfl(hreg, (int)(pos + st - 1));
                if (h <= v[i]) pos += st;
            }
            if (pos == 63 && h63 <= v[i]) pos = 64;  // the lifting steps sum to 63
        } else if (live[i]) {
            uint32_t lo = 0, hi = rng;
            while (lo < hi) {
                const uint32_t m = (lo + hi) >> 1;
                if (head[s_lo + m] <= v[i]) lo = m + 1; else hi = m;
            }
            pos = lo;
        }
        s[i] = (live[i] && pos > 0) ? s_lo + pos - 1 : kNone;
    }

    bool found[V];
#pragma unroll
    for (int i = 0; i < V; ++i) found[i] = false;

    if (CODEC != EXC_PARAPFD && A.tpe == 1) {
        // ---- one lane per candidate: random-access search in its segment
#pragma unroll
        for (int i = 0; i < V; ++i)
            if (s[i] != kNone) found[i] = ra_search<CODEC, SEG>(A, R, s[i], v[i]);
    } else {
        // ---- wave-cooperative: decode each distinct needed segment once
#pragma unroll
        for (int i = 0; i < V; ++i)
            if (live[i]) sarr[ln + 64 * i] = s[i];
        wave_sync();
        uint32_t nd = 0, di[V];
#pragma unroll
        for (int i = 0; i < V; ++i) {
            const uint32_t e = ln + 64 * i;
            const bool first = s[i] != kNone && (e == 0 || sarr[e - 1] != s[i]);
            const uint64_t m = __ballot(first);
            di[i] = nd + below_me(m) + (first ? 1u : 0u) - 1u;
            if (first) dlist[di[i]] = s[i];
            nd += popc(m);
        }
        nd = uni(nd);
        wave_sync();
        for (uint32_t r = 0; r < nd; r += G) {
            const uint32_t gn = min((uint32_t)G, nd - r);
#pragma unroll
            for (int g = 0; g < G; ++g)
                if ((uint32_t)g < gn) decode_seg_block<CODEC, SEG>(A.ix, R, uni(dlist[r + g]), buf + g * SEG);
            wave_sync();
#pragma unroll
            for (int i = 0; i < V; ++i) {
                if (s[i] != kNone && di[i] >= r && di[i] < r + G) {
                    const uint32_t* b = buf + (di[i] - r) * SEG;
                    uint32_t lo = 0, hi = seg_len(R, s[i], SEG);
                    while (lo < hi) {
                        const uint32_t m = (lo + hi) >> 1;
                        if (b[m] < v[i]) lo = m + 1; else hi = m;
                    }
                    found[i] = (lo < seg_len(R, s[i], SEG)) && b[lo] == v[i];
                }
            }
            wave_sync();
        }
    }

    // ---- order-preserving compaction of survivors
    uint32_t cnt = 0, pos[V];
#pragma unroll
    for (int i = 0; i < V; ++i) {
        const uint64_t m = __ballot(found[i]);
        pos[i] = cnt + below_me(m);
        cnt += popc(m);
    }
    wave_sync();
#pragma unroll
    for (int i = 0; i < V; ++i)
        if (found[i]) cand[pos[i]] = v[i];
    wave_sync();
    return uni(cnt);
}

// HS_SegLRC: decode elements [k*SEG, k*SEG+SEG) of list R into dst; returns count.
// Bucket of each element = last bucket start <= its index
// (HS_SegLRC_Intersection_kernel.cu:274-327).
template <int SEG>
__device__ uint32_t decode_hs_chunk(const DevIndex& ix, const TermRec& R, uint32_t k, uint32_t* dst)
{
    constexpr int V = SEG / 64;
    const uint32_t ln = lane_id();
    const uint32_t e0 = k * SEG;
    const uint32_t clen = uni(min((uint32_t)SEG, R.n - e0));
    const uint32_t* starts = ix.hash + R.aux;
    const uint32_t nb = R.nseg;
    const uint32_t b0 = uni(wave_upper_bound(starts, nb + 1, e0) - 1);
    const uint32_t hreg = (b0 + 1 + ln <= nb) ? starts[b0 + 1 + ln] : kNone;
    const uint32_t h63 = uni(__builtin_amdgcn_readlane(hreg, 63));
#pragma unroll
    for (int i = 0; i < V; ++i) {
        const uint32_t j = ln + 64 * i;
        const uint32_t e = e0 + j;
        uint32_t pos = 0;
#pragma unroll
        for (uint32_t st = 32; st >= 1; st >>= 1) {
            const uint32_t h = __shfl(hreg, (int)(pos + st - 1));
            if (h <= e) pos += st;
        }
        if (pos == 63 && h63 <= e) pos = 64;  // the lifting steps sum to 63
        if (j < clen) {
            uint32_t b = b0 + pos;
            if (pos == 64) {
                uint32_t lo = b0 + 65, hi = nb + 1;
                while (lo < hi) {
                    const uint32_t m = (lo + hi) >> 1;
                    if (starts[m] <= e) lo = m + 1; else hi = m;
                }
                b = lo - 1;
            }
            const uint32_t st = starts[b];
            const uint32_t* blk = ix.words + R.woff + ix.seg_off[R.segnum + b];
            const float4 f = ix.fit[R.segnum + b];
            const uint32_t x = e - st;
            dst[j] = bits_at(blk + 1, desc_lb(blk[0]), x) + (uint32_t)(seg_y(x, f.x, f.y) + ix.base[R.segnum + b]);
        }
    }
    return clen;
}

// HS_SegLRC probe: O(1) bucket, fit window, exact fallback
// (HS_SegLRC_Intersection_kernel.cu:334-401 plus the fallback it lacks).
template <int SEG>
__device__ uint32_t probe_hs(const ProbeArgs& A, const TermRec& R, uint32_t* cand, uint32_t nlive)
{
    constexpr int V = SEG / 64;
    const uint32_t ln = lane_id();
    const uint32_t shift = (uint32_t)R.icpt, nb = R.nseg;
    const uint32_t* starts = A.ix.hash + R.aux;
    uint32_t v[V];
    bool found[V];
#pragma unroll
    for (int i = 0; i < V; ++i) {
        const uint32_t e = ln + 64 * i;
        found[i] = false;
        v[i] = 0;
        if (e < nlive) {
            v[i] = cand[e];
            const uint32_t b = v[i] >> shift;
            if (b < nb) {
                const uint32_t st = starts[b], len = starts[b + 1] - st;
                if (len) {
                    const uint32_t* blk = A.ix.words + R.woff + A.ix.seg_off[R.segnum + b];
                    found[i] = fit_block_search(blk, len, A.ix.fit[R.segnum + b], A.ix.base[R.segnum + b], v[i],
                                                A.search != SEARCH_BINARY);
                }
            }
        }
    }
    uint32_t cnt = 0, pos[V];
#pragma unroll
    for (int i = 0; i < V; ++i) {
        const uint64_t m = __ballot(found[i]);
        pos[i] = cnt + below_me(m);
        cnt += popc(m);
    }
    wave_sync();
#pragma unroll
    for (int i = 0; i < V; ++i)
        if (found[i]) cand[pos[i]] = v[i];
    wave_sync();
    return uni(cnt);
}

template <int SEG>
constexpr int wave_lds_words() { return 3 * SEG + 512; }

template <int CODEC, int SEG>
__device__ void run_unit(const ProbeArgs& A, uint32_t u, uint32_t* L)
{
    uint32_t* cand = L;
    uint32_t* sarr = L + SEG;
    uint32_t* dlist = L + 2 * SEG;
    uint32_t* buf = L + 3 * SEG;
    const uint32_t ln = lane_id();
    const uint32_t q = uni(A.unit_q[u]);
    const QueryRec Q = A.queries[q];
    const uint32_t k = u - Q.first_unit;
    const TermRec R0 = A.ix.terms[uni(A.qterms[Q.toff])];
    uint32_t nlive;
    if constexpr (CODEC == EXC_HS_SEGLRC)
        nlive = decode_hs_chunk<SEG>(A.ix, R0, k, cand);
    else
        nlive = decode_seg_block<CODEC, SEG>(A.ix, R0, k, cand);
    nlive = uni(nlive);
    wave_sync();
    for (uint32_t j = 1; j < Q.tc && nlive; ++j) {
        const TermRec R = A.ix.terms[uni(A.qterms[Q.toff + j])];
        if constexpr (CODEC == EXC_HS_SEGLRC)
            nlive = probe_hs<SEG>(A, R, cand, nlive);
        else
            nlive = probe_seg<CODEC, SEG>(A, R, cand, nlive, sarr, dlist, buf);
    }
    uint32_t* out = A.stage + Q.short_off + (uint64_t)k * SEG;
    for (uint32_t e = ln; e < nlive; e += 64) out[e] = cand[e];
    if (ln == 0) A.ucount[u] = nlive;
    wave_sync();
}

template <int CODEC, int SEG>
__global__ __launch_bounds__(256) void k_probe(ProbeArgs A)
{
    __shared__ uint32_t lds[kWaves * wave_lds_words<SEG>()];
    const uint32_t wid = uni(threadIdx.x >> 6);
    uint32_t* L = lds + wid * wave_lds_words<SEG>();
    if (A.persistent) {
        for (;;) {
            uint32_t x = 0;
            if (lane_id() == 0) x = atomicAdd(A.counter, 1u);
            const uint32_t u = uni(__shfl(x, 0));
            if (u >= A.nunits) break;
            run_unit<CODEC, SEG>(A, u, L);
        }
    } else {
        const uint32_t u = blockIdx.x * kWaves + wid;
        if (u < A.nunits) run_unit<CODEC, SEG>(A, u, L);
    }
}

__global__ void k_reset(uint32_t* c) { if (threadIdx.x == 0) *c = 0; }

// ---- scan over unit counts (replaces cudppScan + ucntResult)
__global__ __launch_bounds__(256) void k_scan_local(ScanArgs a)
{
    __shared__ uint32_t wsum[4];
    const uint32_t ln = lane_id(), wid = threadIdx.x >> 6;
    const uint32_t base = blockIdx.x * kScanBlockItems + threadIdx.x * 16;
    uint32_t x[16], tot = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const uint32_t u = base + k;
        x[k] = u < a.nunits ? a.ucount[u] : 0u;
        tot += x[k];
    }
    const uint32_t incl = wave_incl_scan(tot);
    if (ln == 63) wsum[wid] = incl;
    __syncthreads();
    uint32_t woff = 0;
    for (uint32_t w = 0; w < wid; ++w) woff += wsum[w];
    uint32_t run = woff + incl - tot;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const uint32_t u = base + k;
        if (u < a.nunits) a.uoff[u] = run;
        run += x[k];
    }
    if (threadIdx.x == 255) a.bsum[blockIdx.x] = woff + incl;
}

// exclusive scan of the block sums in place; bsum[nblk] = total.
__global__ __launch_bounds__(1024) void k_scan_blocks(uint32_t* bsum, uint32_t nblk)
{
    __shared__ uint32_t wsum[16];
    const uint32_t t = threadIdx.x, ln = lane_id(), wid = t >> 6;
    const uint32_t per = (nblk + 1023) / 1024;
    const uint32_t b0 = t * per;
    uint32_t tot = 0;
    for (uint32_t i = 0; i < per; ++i)
        if (b0 + i < nblk) tot += bsum[b0 + i];
    const uint32_t incl = wave_incl_scan(tot);
    if (ln == 63) wsum[wid] = incl;
    __syncthreads();
    uint32_t woff = 0;
    for (uint32_t w = 0; w < wid; ++w) woff += wsum[w];
    uint32_t run = woff + incl - tot;
    __syncthreads();
    for (uint32_t i = 0; i < per; ++i) {
        if (b0 + i < nblk) {
            const uint32_t x = bsum[b0 + i];
            bsum[b0 + i] = run;
            run += x;
        }
    }
    if (t == 1023) bsum[nblk] = woff + incl;
}

__device__ __forceinline__ uint32_t unit_excl(const ScanArgs& a, uint32_t u)
{
    return u < a.nunits ? a.uoff[u] + a.bsum[u / kScanBlockItems] : a.bsum[(a.nunits + kScanBlockItems - 1) / kScanBlockItems];
}

__global__ __launch_bounds__(256) void k_scatter(ScanArgs a)
{
    const uint32_t ln = lane_id();
    const uint32_t u = blockIdx.x * kWaves + (threadIdx.x >> 6);
    if (u < a.nunits) {
        const uint32_t cnt = a.ucount[u];
        if (cnt) {
            const uint32_t q = a.unit_q[u];
            const QueryRec Q = a.queries[q];
            const uint32_t* src = a.stage + Q.short_off + (uint64_t)(u - Q.first_unit) * a.seg;
            uint32_t* dst = a.result + unit_excl(a, u);
            for (uint32_t e = ln; e < cnt; e += 64) dst[e] = src[e];
        }
    }
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < a.nq) {
        const QueryRec Q = a.queries[g];
        a.qcum[g] = unit_excl(a, Q.first_unit + Q.nunits);
    }
}

// ---- whole-list decompression (the *_Decompression programs)
template <int CODEC, int SEG>
__global__ __launch_bounds__(256) void k_decompress(DevIndex ix, const uint32_t* terms, const uint64_t* out_off,
                                                    const uint32_t* unit_term, const uint32_t* unit_k,
                                                    uint32_t nunits, uint32_t* out)
{
    __shared__ uint32_t lds[kWaves * SEG];
    const uint32_t wid = uni(threadIdx.x >> 6), ln = lane_id();
    const uint32_t u = blockIdx.x * kWaves + wid;
    if (u >= nunits) return;
    uint32_t* L = lds + wid * SEG;
    const uint32_t ti = uni(unit_term[u]), k = uni(unit_k[u]);
    const TermRec R = ix.terms[uni(terms[ti])];
    uint32_t n;
    if constexpr (CODEC == EXC_HS_SEGLRC)
        n = decode_hs_chunk<SEG>(ix, R, k, L);
    else
        n = decode_seg_block<CODEC, SEG>(ix, R, k, L);
    wave_sync();
    uint32_t* o = out + out_off[ti] + (uint64_t)k * SEG;
    for (uint32_t e = ln; e < n; e += 64) o[e] = L[e];
}

template <int CODEC>
hipError_t probe_for_seg(const ProbeArgs& a, uint32_t grid, hipStream_t s)
{
    if (a.ix.seg == 256)
        hipLaunchKernelGGL((k_probe<CODEC, 256>), dim3(grid), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((k_probe<CODEC, 128>), dim3(grid), dim3(256), 0, s, a);
    return hipGetLastError();
}

template <int CODEC>
hipError_t decomp_for_seg(const DevIndex& ix, const uint32_t* terms, const uint64_t* out_off,
                          const uint32_t* ut, const uint32_t* uk, uint32_t n, uint32_t* out, hipStream_t s)
{
    const uint32_t grid = (n + kWaves - 1) / kWaves;
    if (ix.seg == 256)
        hipLaunchKernelGGL((k_decompress<CODEC, 256>), dim3(grid), dim3(256), 0, s, ix, terms, out_off, ut, uk, n, out);
    else
        hipLaunchKernelGGL((k_decompress<CODEC, 128>), dim3(grid), dim3(256), 0, s, ix, terms, out_off, ut, uk, n, out);
    return hipGetLastError();
}

}  // namespace

uint32_t probe_waves_per_block() { return kWaves; }

template <int CODEC>
static int occ_for(uint32_t seg)
{
    int nb = 0;
    hipError_t e = (seg == 256)
        ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(&k_probe<CODEC, 256>), 256, 0)
        : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(&k_probe<CODEC, 128>), 256, 0);
    return (e == hipSuccess && nb > 0) ? nb : 4;
}

int probe_blocks_per_cu(int codec, uint32_t seg)
{
    switch (codec) {
    case EXC_LRC: return occ_for<EXC_LRC>(seg);
    case EXC_SEGLRC: return occ_for<EXC_SEGLRC>(seg);
    case EXC_HS_SEGLRC: return occ_for<EXC_HS_SEGLRC>(seg);
    case EXC_PARAPFD: return occ_for<EXC_PARAPFD>(seg);
    }
    return 4;
}

hipError_t launch_reset(uint32_t* counter, hipStream_t s)
{
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, counter);
    return hipGetLastError();
}

hipError_t launch_probe(const ProbeArgs& a, uint32_t grid, hipStream_t s)
{
    switch (a.ix.codec) {
    case EXC_LRC: return probe_for_seg<EXC_LRC>(a, grid, s);
    case EXC_SEGLRC: return probe_for_seg<EXC_SEGLRC>(a, grid, s);
    case EXC_HS_SEGLRC: return probe_for_seg<EXC_HS_SEGLRC>(a, grid, s);
    case EXC_PARAPFD: return probe_for_seg<EXC_PARAPFD>(a, grid, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_scan(const ScanArgs& a, hipStream_t s)
{
    const uint32_t nblk = (a.nunits + kScanBlockItems - 1) / kScanBlockItems;
    if (nblk) hipLaunchKernelGGL(k_scan_local, dim3(nblk), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, s, a.bsum, nblk);
    const uint32_t g1 = (a.nunits + kWaves - 1) / kWaves, g2 = (a.nq + 255) / 256;
    const uint32_t grid = g1 > g2 ? g1 : g2;
    if (grid) hipLaunchKernelGGL(k_scatter, dim3(grid), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_decompress(const DevIndex& ix, const uint32_t* terms, const uint64_t* out_off,
                             const uint32_t* ut, const uint32_t* uk, uint32_t n, uint32_t* out, hipStream_t s)
{
    if (!n) return hipSuccess;
    switch (ix.codec) {
    case EXC_LRC: return decomp_for_seg<EXC_LRC>(ix, terms, out_off, ut, uk, n, out, s);
    case EXC_SEGLRC: return decomp_for_seg<EXC_SEGLRC>(ix, terms, out_off, ut, uk, n, out, s);
    case EXC_HS_SEGLRC: return decomp_for_seg<EXC_HS_SEGLRC>(ix, terms, out_off, ut, uk, n, out, s);
    case EXC_PARAPFD: return decomp_for_seg<EXC_PARAPFD>(ix, terms, out_off, ut, uk, n, out, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace exc
