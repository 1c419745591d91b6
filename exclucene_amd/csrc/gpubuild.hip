// gpubuild.hip -- index construction on the GPU: seeded synthetic posting
// lists and the SegLRC, LRC and ParaPFD encoders, so that an index of tens of
// GB is generated and encoded where it is served instead of crossing PCIe.
//
// The encoder follows the host builder block for block: the per-segment fit,
// base and bit width come from exc_encode.h (the same code, the same IEEE
// operation order: both sides build with -ffp-contract=off), the ParaPFor
// block is a descriptor word (lb << 26, no exceptions: SegLRC uses FRAC 0) and
// the residuals packed LSB first (GI/ParaPFor.cpp:56-126).  Segment-level bytes
// are therefore identical to the host builder's; only the list-level fit of
// .regression_info_list sums in a different order (a block reduction).
#include "exc_encode.h"
#include "exc_format.h"
#include "exc_internal.h"
#include "exc_model.h"

namespace exc {
namespace {

// ---------------------------------------------------------------- synthesis

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// One lane's share [lo, hi) of an item's docID span.  Dense lists (p >= 1/32)
// keep each docID whose hash falls under p * 2^32; sparse lists draw geometric
// gaps (a Bernoulli(p) process, so disjoint spans are independent).
template <bool WRITE>
__device__ uint32_t synth_lane(const SynthArgs& a, uint32_t t, uint32_t c, uint32_t ln, uint64_t lo, uint64_t hi,
                               uint32_t* out)
{
    const double p = (double)a.df[t] / (double)a.universe;
    uint32_t n = 0;
    if (p >= 1.0 / 32.0) {
        const uint64_t th = p >= 1.0 ? 0x100000000ull : (uint64_t)(p * 4294967296.0);
        const uint64_t key = a.seed ^ ((uint64_t)t * 0x9E3779B97F4A7C15ull);
        for (uint64_t d = lo; d < hi; ++d) {
            if ((mix64(key + d * 0xD1B54A32D192ED03ull) >> 32) < th) {
                if (WRITE) out[n] = (uint32_t)d;
                ++n;
            }
        }
    } else if (p > 0.0) {
        uint64_t s = mix64(a.seed ^ ((uint64_t)t * 0x9E3779B97F4A7C15ull) ^ ((uint64_t)c * 0xC2B2AE3D27D4EB4Full) ^
                           ((uint64_t)ln * 0x165667B19E3779F9ull));
        const double lq = log1p(-p);
        uint64_t d = lo;
        for (;;) {
            s += 0x9E3779B97F4A7C15ull;
            const double u = (double)((mix64(s) >> 11) + 1) * 0x1.0p-53;  // (0, 1]
            const double g = floor(log(u) / lq);
            if (g >= (double)(hi - d)) break;
            d += (uint64_t)g;
            if (WRITE) out[n] = (uint32_t)d;
            ++n;
            ++d;
            if (d >= hi) break;
        }
    }
    return n;
}

// One wave per item (term t, chunk c of its docID range).  Counting pass:
// counts[i] = docIDs of the item.  Writing pass: each lane counts its share,
// a wave scan places it, and the lane writes its docIDs in order.
template <bool WRITE>
__global__ __launch_bounds__(256) void k_synth(SynthArgs a)
{
    const uint32_t wid = threadIdx.x >> 6, ln = threadIdx.x & 63u;
    const uint64_t i = (uint64_t)blockIdx.x * 4 + wid + a.item0;
    if (i >= a.item1) return;
    const uint32_t t = a.item_term[i], c = a.item_chunk[i], nc = a.nchunks[t];
    const uint64_t lo = (uint64_t)c * a.universe / nc, hi = (uint64_t)(c + 1) * a.universe / nc;
    const uint64_t span = hi - lo;
    const uint64_t llo = lo + span * ln / 64, lhi = lo + span * (ln + 1) / 64;
    const uint32_t cnt = synth_lane<false>(a, t, c, ln, llo, lhi, nullptr);
    // inclusive wave scan of the lane counts
    uint32_t incl = cnt;
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(incl, d);
        if (ln >= d) incl += y;
    }
    if (!WRITE) {
        if (ln == 63) a.counts[i] = incl;
        return;
    }
    uint32_t* out = a.docs + (a.item_off[i] - a.doc_base) + (incl - cnt);
    synth_lane<true>(a, t, c, ln, llo, lhi, out);
}

// ---------------------------------------------------------------- encoder

__device__ __forceinline__ uint32_t term_of_segment(const uint32_t* seg_num, uint32_t nterms, uint32_t g)
{
    // last t with seg_num[t] <= g (terms without segments share the next one's start)
    uint32_t lo = 0, hi = nterms;
    while (hi - lo > 1) {
        const uint32_t m = (lo + hi) >> 1;
        if (seg_num[m] <= g) lo = m; else hi = m;
    }
    return lo;
}

// Thread per segment: fit, base, bit width, head; the block's word count.
__global__ __launch_bounds__(256) void k_seg_fit(EncodeArgs a)
{
    const uint64_t g64 = (uint64_t)blockIdx.x * 256 + threadIdx.x + a.g0;
    if (g64 >= a.g1) return;
    const uint32_t g = (uint32_t)g64;
    const uint32_t t = term_of_segment(a.seg_num, a.nterms, g);
    const uint32_t k = g - a.seg_num[t];
    const uint64_t nt = a.list_off[t + 1] - a.list_off[t];
    const uint32_t st = k * a.seg, n = (uint32_t)min((uint64_t)a.seg, nt - st);
    const uint32_t* L = a.docs + (a.list_off[t] - a.doc_base) + st;
    const exc_seg_info f = seg_fit(L, n);
    int32_t base;
    uint32_t lb;
    seg_base_bits(L, n, f, base, lb);
    SegRec q;
    q.off = 0;
    q.desc = lb << 26;
    q.slope = f.fSlope;
    q.icpt = f.fIntercept;
    q.rl = f.nRangeLeft;
    q.rr = f.nRangeRight;
    q.base = base;
    q.head = L[0];
    if (!(q.slope > 0.0f)) {
        q.slope = 1.0f;
        q.rl = q.rr = 0x7fffffffu;
    }
    a.rec[g] = q;
    a.head[g] = L[0];
    a.nwords[g] = 1 + words_for(n, lb);
}

// Thread per segment: descriptor and packed residuals at the block's global
// word offset; the record's offset relative to its list.
__global__ __launch_bounds__(256) void k_seg_pack(EncodeArgs a)
{
    const uint64_t g64 = (uint64_t)blockIdx.x * 256 + threadIdx.x + a.g0;
    if (g64 >= a.g1) return;
    const uint32_t g = (uint32_t)g64;
    const uint32_t t = term_of_segment(a.seg_num, a.nterms, g);
    const uint32_t k = g - a.seg_num[t];
    const uint64_t nt = a.list_off[t + 1] - a.list_off[t];
    const uint32_t st = k * a.seg, n = (uint32_t)min((uint64_t)a.seg, nt - st);
    const uint32_t* L = a.docs + (a.list_off[t] - a.doc_base) + st;
    SegRec q = a.rec[g];
    const uint64_t go = a.goff[g];
    q.off = (uint32_t)(go - a.goff[a.seg_num[t]]);
    a.rec[g].off = q.off;
    const uint32_t lb = desc_lb(q.desc);
    exc_seg_info f{q.slope, q.icpt, 0, 0};
    uint32_t* o = a.words + go;
    o[0] = q.desc;
    ++o;
    uint64_t acc = 0;
    uint32_t nb = 0;
    for (uint32_t j = 0; j < n; ++j) {
        const uint32_t v = (uint32_t)(seg_residual(L, j, f) - q.base);
        acc |= (uint64_t)v << nb;
        nb += lb;
        if (nb >= 32) {
            *o++ = (uint32_t)acc;
            acc >>= 32;
            nb -= 32;
        }
    }
    if (nb) *o = (uint32_t)acc;
}

// Block per list: the list-level fit of .regression_info_list (double sums as
// a block reduction, float slope and intercept, ranges = max deviations + 1).
__global__ __launch_bounds__(256) void k_list_fit(EncodeArgs a)
{
    __shared__ double r0[256], r1[256];
    __shared__ float m0[256], m1[256];
    const uint32_t t = a.t0 + blockIdx.x;
    if (t >= a.t1) return;
    TermRec& R = a.terms[t];
    const uint64_t n = a.list_off[t + 1] - a.list_off[t];
    const uint32_t* L = a.docs + (a.list_off[t] - a.doc_base);
    const uint32_t tid = threadIdx.x;
    if (n < 2) {
        if (tid == 0) {
            R.slope = 1.0f;
            R.range_l = R.range_r = 4.0e9f;
            const float ic = n ? (float)L[0] : 0.0f;
            R.icpt = __float_as_int(ic);
        }
        return;
    }
    double s = 0;
    for (uint64_t i = tid; i < n; i += 256) s += (double)L[i];
    r0[tid] = s;
    __syncthreads();
    for (uint32_t w = 128; w; w >>= 1) {
        if (tid < w) r0[tid] += r0[tid + w];
        __syncthreads();
    }
    const double ya = r0[0] / (double)n, xa = (double)(n - 1) * 0.5;
    __syncthreads();
    double sx = 0, sxy = 0;
    for (uint64_t i = tid; i < n; i += 256) {
        const double dx = (double)i - xa;
        sx += dx * dx;
        sxy += dx * ((double)L[i] - ya);
    }
    r0[tid] = sx;
    r1[tid] = sxy;
    __syncthreads();
    for (uint32_t w = 128; w; w >>= 1) {
        if (tid < w) {
            r0[tid] += r0[tid + w];
            r1[tid] += r1[tid + w];
        }
        __syncthreads();
    }
    const float slope = (float)(r1[0] / r0[0]);
    const float icpt = (float)(ya - (double)slope * xa);
    float rl = 0, rr = 0;
    for (uint64_t i = tid; i < n; i += 256) {
        const float d = ((float)L[i] - icpt) / slope - (float)i;
        rl = fmaxf(rl, d);
        rr = fminf(rr, d);
    }
    m0[tid] = rl;
    m1[tid] = rr;
    __syncthreads();
    for (uint32_t w = 128; w; w >>= 1) {
        if (tid < w) {
            m0[tid] = fmaxf(m0[tid], m0[tid + w]);
            m1[tid] = fminf(m1[tid], m1[tid + w]);
        }
        __syncthreads();
    }
    if (tid == 0) {
        if (slope > 0.0f) {
            R.slope = slope;
            R.icpt = __float_as_int(icpt);
            R.range_l = (float)((uint32_t)m0[0] + 1u);
            R.range_r = (float)((uint32_t)(0.0f - m1[0]) + 1u);
        } else {
            R.slope = 1.0f;
            R.range_l = R.range_r = 4.0e9f;
        }
    }
}

// ---------------------------------------------------------------- LRC and ParaPFD

// Thread per list: the LRC list fit and the residuals' minimum folded into the
// intercept (the host builder's code, exc_encode.h), kept in the term record
// for the block passes.
__global__ __launch_bounds__(256) void k_lrc_lists(EncodeArgs a)
{
    const uint32_t t = a.t0 + blockIdx.x * 256 + threadIdx.x;
    if (t >= a.t1) return;
    const uint32_t n = (uint32_t)(a.list_off[t + 1] - a.list_off[t]);
    const uint32_t* L = a.docs + (a.list_off[t] - a.doc_base);
    exc_lrc_info f = lrc_fit(L, n);
    f.iIntercept = (int32_t)((uint32_t)f.iIntercept + (uint32_t)lrc_min_residual(L, n, f));
    TermRec& R = a.terms[t];
    R.slope = f.fSlope;
    R.icpt = f.iIntercept;
    R.range_l = R.range_r = f.fSearchRange;
}

// The values a block of an LRC or ParaPFD list stores: residuals against the
// list fit (GC/LRC/LRC_Generator/LRC_Generator_kernel.cu:68-88) or d-gaps, the
// list's first docID standing for its own gap
// (GC/ParaPFD/ParaPFD_Compression/main.cpp:163-185).
struct BlockValues {
    const uint32_t* L;  // the list
    uint32_t st;        // the block's first index
    int32_t codec;
    float slope;
    int32_t icpt;
    __device__ uint32_t operator()(uint32_t i) const
    {
        const uint32_t p = st + i;
        if (codec == EXC_LRC) return L[p] - lrc_pred(p, slope, icpt);
        return p ? L[p] - L[p - 1] : L[0];
    }
};

__device__ __forceinline__ BlockValues block_values(const EncodeArgs& a, uint32_t t, uint32_t k, uint32_t& n)
{
    const uint64_t nt = a.list_off[t + 1] - a.list_off[t];
    BlockValues v;
    v.L = a.docs + (a.list_off[t] - a.doc_base);
    v.st = k * a.seg;
    v.codec = a.codec;
    v.slope = a.codec == EXC_LRC ? a.terms[t].slope : 1.0f;
    v.icpt = a.codec == EXC_LRC ? a.terms[t].icpt : 0;
    n = (uint32_t)min((uint64_t)a.seg, nt - v.st);
    return v;
}

// Thread per block: descriptor and size by the host builder's ParaPFor rule
// (exc_encode.h), head and (ParaPFD) median -- the .segMedian row, kept in
// the record's rl field.
__global__ __launch_bounds__(256) void k_blk_fit(EncodeArgs a)
{
    const uint64_t g64 = (uint64_t)blockIdx.x * 256 + threadIdx.x + a.g0;
    if (g64 >= a.g1) return;
    const uint32_t g = (uint32_t)g64;
    const uint32_t t = term_of_segment(a.seg_num, a.nterms, g);
    uint32_t n;
    const BlockValues val = block_values(a, t, g - a.seg_num[t], n);
    PforPlan p = pfor_plan(val, n, a.frac);
    if (!p.words) {
        atomicOr(a.err, 1u);
        p.desc = 0;
        p.words = 1;
    }
    SegRec q;
    q.off = 0;
    q.desc = p.desc;
    q.slope = 1.0f;
    q.icpt = 0.0f;
    q.rl = a.codec == EXC_PARAPFD ? val.L[val.st + n / 2] : 0u;
    // ParaPFD: the values at the other two quarter positions
    q.rr = a.codec == EXC_PARAPFD ? val.L[val.st + pfd_m3(n)] : 0u;
    q.base = a.codec == EXC_PARAPFD ? (int32_t)val.L[val.st + pfd_m1(n)] : 0;
    if (a.codec == EXC_PARAPFD && a.pfd_oct)
        for (uint32_t k = 0; k < 4; ++k) a.pfd_oct[4ull * g + k] = val.L[val.st + pfd_oct_pos(n, k)];
    q.head = val.L[val.st];
    a.rec[g] = q;
    a.head[g] = q.head;
    a.nwords[g] = p.words;
}

// Thread per block: its words at the block's global offset; the record's
// offset relative to its list.
__global__ __launch_bounds__(256) void k_blk_pack(EncodeArgs a)
{
    const uint64_t g64 = (uint64_t)blockIdx.x * 256 + threadIdx.x + a.g0;
    if (g64 >= a.g1) return;
    const uint32_t g = (uint32_t)g64;
    const uint32_t t = term_of_segment(a.seg_num, a.nterms, g);
    uint32_t n;
    const BlockValues val = block_values(a, t, g - a.seg_num[t], n);
    const uint64_t go = a.goff[g];
    a.rec[g].off = (uint32_t)(go - a.goff[a.seg_num[t]]);
    pfor_write(val, n, a.rec[g].desc, a.words + go);
}

// ---------------------------------------------------------------- HS_SegLRC

// Thread per list: the bucket shift and count of its hash table, by the host
// builder's rule (GC/HS_SegLRC/HS_SegLRC_Generator/HS_SegLRC_Generator.cu:503-594):
// a fixed shift, or bits(max) - bits(ceil(n / target)), and one bucket
// (shift = bits(max)) for lists of at most `target` docIDs.
__global__ __launch_bounds__(256) void k_hs_lists(EncodeArgs a)
{
    const uint32_t t = a.t0 + blockIdx.x * 256 + threadIdx.x;
    if (t >= a.t1) return;
    const uint32_t n = (uint32_t)(a.list_off[t + 1] - a.list_off[t]);
    const uint32_t* L = a.docs + (a.list_off[t] - a.doc_base);
    const uint32_t maxb = n ? bits_of(L[n - 1]) : 0;
    uint32_t shift = a.hs_fixed;
    if (!shift) shift = n <= a.hs_target ? maxb : maxb - bits_of((n + a.hs_target - 1) / a.hs_target);
    a.hs_shift[t] = shift;
    a.hs_nb[t] = n ? (L[n - 1] >> shift) + 1 : 1;
}

// first index of L[0 .. n) whose docID is >= v
__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t* L, uint32_t n, uint64_t v)
{
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t m = (lo + hi) >> 1;
        if ((uint64_t)L[m] < v) lo = m + 1; else hi = m;
    }
    return lo;
}

// bucket g of list t: [start, end) of the list
__device__ __forceinline__ void hs_bucket(const EncodeArgs& a, uint32_t t, uint32_t b, const uint32_t* L, uint32_t n,
                                          uint32_t& start, uint32_t& end)
{
    const uint32_t shift = a.hs_shift[t], nb = a.hs_nb[t];
    start = b ? lower_bound_u32(L, n, (uint64_t)b << shift) : 0u;
    end = b + 1 < nb ? lower_bound_u32(L, n, (uint64_t)(b + 1) << shift) : n;
}

// Thread per bucket: its table entry (bucket 0 writes the header, the last
// bucket the closing n), the SegLRC fit and base of its docIDs (exc_encode.h,
// the host builder's code) and its block size; an empty bucket has no block.
__global__ __launch_bounds__(256) void k_hs_fit(EncodeArgs a)
{
    const uint64_t g64 = (uint64_t)blockIdx.x * 256 + threadIdx.x + a.g0;
    if (g64 >= a.g1) return;
    const uint32_t g = (uint32_t)g64;
    const uint32_t t = term_of_segment(a.seg_num, a.nterms, g);
    const uint32_t b = g - a.seg_num[t];
    const uint32_t n = (uint32_t)(a.list_off[t + 1] - a.list_off[t]);
    const uint32_t* L = a.docs + (a.list_off[t] - a.doc_base);
    uint32_t start, end;
    hs_bucket(a, t, b, L, n, start, end);
    uint32_t* tab = a.hash + a.hs_off[t];
    if (b == 0) {
        tab[0] = a.hs_shift[t];
        tab[1] = a.hs_nb[t];
    }
    tab[2 + b] = start;
    if (b + 1 == a.hs_nb[t]) tab[3 + b] = n;
    const uint32_t len = end - start;
    const exc_seg_info f = seg_fit(L + start, len);
    int32_t base;
    uint32_t lb;
    seg_base_bits(L + start, len, f, base, lb);
    SegRec q;
    q.off = 0;
    q.desc = lb << 26;
    q.slope = f.fSlope;
    q.icpt = f.fIntercept;
    q.rl = f.nRangeLeft;
    q.rr = f.nRangeRight;
    q.base = base;
    q.head = start;
    a.rec[g] = q;
    a.nwords[g] = len ? 1 + words_for(len, lb) : 0u;
}

// Thread per bucket: descriptor and packed residuals (none for an empty bucket,
// whose offset is the next block's, as in the reference's files).
__global__ __launch_bounds__(256) void k_hs_pack(EncodeArgs a)
{
    const uint64_t g64 = (uint64_t)blockIdx.x * 256 + threadIdx.x + a.g0;
    if (g64 >= a.g1) return;
    const uint32_t g = (uint32_t)g64;
    const uint32_t t = term_of_segment(a.seg_num, a.nterms, g);
    const uint32_t b = g - a.seg_num[t];
    const uint32_t n = (uint32_t)(a.list_off[t + 1] - a.list_off[t]);
    const uint32_t* L = a.docs + (a.list_off[t] - a.doc_base);
    const SegRec q = a.rec[g];
    const uint64_t go = a.goff[g];
    a.rec[g].off = (uint32_t)(go - a.goff[a.seg_num[t]]);
    const uint32_t* tab = a.hash + a.hs_off[t];
    const uint32_t start = tab[2 + b], len = tab[3 + b] - start;
    (void)n;
    if (!len) return;
    const exc_seg_info f{q.slope, q.icpt, 0, 0};
    const uint32_t lb = desc_lb(q.desc);
    uint32_t* o = a.words + go;
    o[0] = q.desc;
    BitWriter w(o + 1);
    for (uint32_t j = 0; j < len; ++j) w.put((uint32_t)(seg_residual(L + start, j, f) - q.base), lb);
    w.flush();
}

// Thread per term: the record fields that do not depend on the docIDs.
__global__ __launch_bounds__(256) void k_term_recs(EncodeArgs a)
{
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= a.nterms) return;
    TermRec R;
    R.woff = a.goff[a.seg_num[t]];
    R.n = (uint32_t)(a.list_off[t + 1] - a.list_off[t]);
    R.segnum = a.seg_num[t];
    R.nseg = a.seg_num[t + 1] - a.seg_num[t];
    R.aux = 0;
    R.slope = 1.0f;
    R.icpt = 0;
    R.range_l = R.range_r = 4.0e9f;
    if (a.codec == EXC_HS_SEGLRC) {  // buckets: count, table position, shift
        R.aux = (uint32_t)(a.hs_off[t] + 2);
        R.icpt = (int32_t)a.hs_shift[t];
    }
    if (a.codec == EXC_LRC) {  // the list fit of k_lrc_lists, as the upload keeps it
        const TermRec& F = a.terms[t];
        R.icpt = F.icpt;
        if (F.slope > 0.0f) {
            R.slope = F.slope;
            R.range_l = R.range_r = F.range_l;
        }
    }
    R.model = kNoModel;
    R.term = t;
    R.merr = R.mx0 = R.mxs = 0.0f;
    R.mkind = 0;
    a.terms[t] = R;
}

// Thread per ParaPFD list: the list-level line through its block heads, as
// the upload derives it (heads_line).
__global__ __launch_bounds__(256) void k_heads_line(EncodeArgs a)
{
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= a.nterms) return;
    TermRec& R = a.terms[t];
    float icpt;
    heads_line(a.head + a.seg_num[t], a.seg_num[t + 1] - a.seg_num[t], a.seg, R.slope, icpt, R.range_l, R.range_r);
    R.icpt = __float_as_int(icpt);
}

// Thread per ParaPFD block of a decoded run: the values at its quarter
// positions into its record (base: pfd_m1, rr: pfd_m3) and those at the
// quarters' midpoints into oct.
__global__ __launch_bounds__(256) void k_pfd_anchors(const AnchorJob* jobs, uint32_t n, const uint32_t* dec, SegRec* rec,
                                                     uint32_t* oct)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const AnchorJob j = jobs[i];
    rec[j.rec].base = (int32_t)dec[j.pos + pfd_m1(j.len)];
    rec[j.rec].rr = dec[j.pos + pfd_m3(j.len)];
    for (uint32_t k = 0; k < 4; ++k) oct[4 * j.rec + k] = dec[j.pos + pfd_oct_pos(j.len, k)];
}

// ---------------------------------------------------------------- HS_SegLRC compaction

// Thread per bucket (and one past the last): 1 if the bucket holds docIDs.
__global__ __launch_bounds__(256) void k_hsc_flags(HsCompactArgs a)
{
    const uint64_t g64 = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (g64 > a.nbuckets) return;
    if (g64 == a.nbuckets) {
        a.flag[g64] = 0;
        return;
    }
    const uint32_t g = (uint32_t)g64;
    const uint32_t t = term_of_segment(a.seg_num, a.nterms, g);
    const uint32_t b = g - a.seg_num[t];
    const uint32_t* tab = a.hash + a.hs_off[t];
    a.flag[g] = tab[3 + b] > tab[2 + b] ? 1u : 0u;
}

// Thread per list: its records (non-empty buckets).
__global__ __launch_bounds__(256) void k_hsc_counts(HsCompactArgs a)
{
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= a.nterms) return;
    a.count[t] = (uint32_t)(a.cidx[a.seg_num[t + 1]] - a.cidx[a.seg_num[t]]);
}

// Thread per bucket: a non-empty bucket's record and first element to its place.
__global__ __launch_bounds__(256) void k_hsc_recs(HsCompactArgs a)
{
    const uint64_t g = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= a.nbuckets || !a.flag[g]) return;
    const SegRec q = a.rec[g];
    const uint64_t c = a.cidx[g];
    a.crec[c] = q;
    a.chead[c] = q.head;
}

// Thread per occupancy group (32 buckets of a list with empty buckets): its
// bit word and the records before it; group 0 also writes the header.
__global__ __launch_bounds__(256) void k_hsc_occ(HsCompactArgs a)
{
    const uint64_t j = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= a.ngroups) return;
    uint32_t lo = 0, hi = a.nterms;  // last list whose groups start at or before j
    while (hi - lo > 1) {
        const uint32_t m = (lo + hi) >> 1;
        if (a.grp_off[m] <= j) lo = m; else hi = m;
    }
    const uint32_t t = lo;
    const uint32_t grp = (uint32_t)(j - a.grp_off[t]);
    const uint32_t nb = a.hash[a.hs_off[t] + 1];
    const uint64_t g0 = (uint64_t)a.seg_num[t] + 32ull * grp;
    uint32_t bits = 0;
    for (uint32_t i = 0; i < 32 && 32 * grp + i < nb; ++i) bits |= a.flag[g0 + i] << i;
    uint32_t* o = a.occ + a.occ_off[t];
    if (grp == 0) {
        o[0] = nb;
        o[1] = 0;
    }
    o[2 + 2 * grp] = bits;
    o[3 + 2 * grp] = (uint32_t)(a.cidx[g0] - a.cidx[a.seg_num[t]]);
}

// Thread per list: the record fields of the compact layout.
__global__ __launch_bounds__(256) void k_hsc_terms(HsCompactArgs a)
{
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= a.nterms) return;
    TermRec R = a.terms[t];
    R.segnum = (uint32_t)a.cidx[a.seg_num[t]];
    R.nseg = a.count[t];
    R.aux = a.occ_off[t] == ~0ull ? kDenseBuckets : (uint32_t)a.occ_off[t];
    a.terms[t] = R;
}

// ---------------------------------------------------------------- scan (u64)

constexpr uint32_t kScanItems = 4096;  // 256 threads x 16

template <class T>
__global__ __launch_bounds__(256) void k_block_scan(const T* in, uint64_t* out, uint64_t* bsum, uint64_t n)
{
    __shared__ uint64_t ws[4];
    const uint32_t tid = threadIdx.x, ln = tid & 63u, wid = tid >> 6;
    const uint64_t base = (uint64_t)blockIdx.x * kScanItems + (uint64_t)tid * 16;
    uint64_t x[16], tot = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        x[k] = base + k < n ? (uint64_t)in[base + k] : 0ull;
        tot += x[k];
    }
    uint64_t incl = tot;
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint64_t y = __shfl_up(incl, d);
        if (ln >= d) incl += y;
    }
    if (ln == 63) ws[wid] = incl;
    __syncthreads();
    uint64_t woff = 0;
    for (uint32_t w = 0; w < wid; ++w) woff += ws[w];
    uint64_t run = woff + incl - tot;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        if (base + k < n) out[base + k] = run;
        run += x[k];
    }
    if (tid == 255) bsum[blockIdx.x] = woff + incl;
}

__global__ __launch_bounds__(256) void k_add_offsets(uint64_t* out, const uint64_t* boff, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] += boff[i / kScanItems];
}

template <class T>
hipError_t scan_rec(const T* in, uint64_t* out, uint64_t n, hipStream_t s)
{
    if (!n) return hipSuccess;
    const uint64_t nb = (n + kScanItems - 1) / kScanItems;
    uint64_t* bsum = nullptr;
    uint64_t* boff = nullptr;
    hipError_t e = hipMallocAsync((void**)&bsum, nb * 8, s);
    if (e == hipSuccess) e = hipMallocAsync((void**)&boff, nb * 8, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_block_scan<T>, dim3((uint32_t)nb), dim3(256), 0, s, in, out, bsum, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess && nb > 1) e = scan_rec<uint64_t>(bsum, boff, nb, s);
    if (e == hipSuccess && nb > 1) {
        hipLaunchKernelGGL(k_add_offsets, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, out, boff, n);
        e = hipGetLastError();
    }
    if (bsum) (void)hipFreeAsync(bsum, s);
    if (boff) (void)hipFreeAsync(boff, s);
    return e;
}

}  // namespace

hipError_t launch_synth(const SynthArgs& a, bool write, hipStream_t s)
{
    if (a.item1 <= a.item0) return hipSuccess;
    const uint64_t grid = (a.item1 - a.item0 + 3) / 4;
    if (write)
        hipLaunchKernelGGL(k_synth<true>, dim3((uint32_t)grid), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(k_synth<false>, dim3((uint32_t)grid), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_seg_fit(const EncodeArgs& a, hipStream_t s)
{
    if (a.g1 <= a.g0) return hipSuccess;
    hipLaunchKernelGGL(k_seg_fit, dim3((uint32_t)((a.g1 - a.g0 + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_seg_pack(const EncodeArgs& a, hipStream_t s)
{
    if (a.g1 > a.g0)
        hipLaunchKernelGGL(k_seg_pack, dim3((uint32_t)((a.g1 - a.g0 + 255) / 256)), dim3(256), 0, s, a);
    if (a.t1 > a.t0) hipLaunchKernelGGL(k_list_fit, dim3(a.t1 - a.t0), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_lrc_lists(const EncodeArgs& a, hipStream_t s)
{
    if (a.t1 > a.t0) hipLaunchKernelGGL(k_lrc_lists, dim3((a.t1 - a.t0 + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_blk_fit(const EncodeArgs& a, hipStream_t s)
{
    if (a.g1 > a.g0) hipLaunchKernelGGL(k_blk_fit, dim3((uint32_t)((a.g1 - a.g0 + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_blk_pack(const EncodeArgs& a, hipStream_t s)
{
    if (a.g1 > a.g0) hipLaunchKernelGGL(k_blk_pack, dim3((uint32_t)((a.g1 - a.g0 + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_hs_lists(const EncodeArgs& a, hipStream_t s)
{
    if (a.t1 > a.t0) hipLaunchKernelGGL(k_hs_lists, dim3((a.t1 - a.t0 + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_hs_fit(const EncodeArgs& a, hipStream_t s)
{
    if (a.g1 > a.g0) hipLaunchKernelGGL(k_hs_fit, dim3((uint32_t)((a.g1 - a.g0 + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_hs_pack(const EncodeArgs& a, hipStream_t s)
{
    if (a.g1 > a.g0) hipLaunchKernelGGL(k_hs_pack, dim3((uint32_t)((a.g1 - a.g0 + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_term_recs(const EncodeArgs& a, hipStream_t s)
{
    if (!a.nterms) return hipSuccess;
    hipLaunchKernelGGL(k_term_recs, dim3((a.nterms + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_heads_line(const EncodeArgs& a, hipStream_t s)
{
    if (!a.nterms) return hipSuccess;
    hipLaunchKernelGGL(k_heads_line, dim3((a.nterms + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_pfd_anchors(const AnchorJob* jobs, uint32_t n, const uint32_t* dec, SegRec* rec, uint32_t* oct,
                              hipStream_t s)
{
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_pfd_anchors, dim3((n + 255) / 256), dim3(256), 0, s, jobs, n, dec, rec, oct);
    return hipGetLastError();
}

hipError_t launch_hs_compact_flags(const HsCompactArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(k_hsc_flags, dim3((uint32_t)((a.nbuckets + 256) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_hs_compact_counts(const HsCompactArgs& a, hipStream_t s)
{
    if (a.nterms) hipLaunchKernelGGL(k_hsc_counts, dim3((a.nterms + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_hs_compact(const HsCompactArgs& a, hipStream_t s)
{
    if (a.nbuckets)
        hipLaunchKernelGGL(k_hsc_recs, dim3((uint32_t)((a.nbuckets + 255) / 256)), dim3(256), 0, s, a);
    if (a.ngroups) hipLaunchKernelGGL(k_hsc_occ, dim3((uint32_t)((a.ngroups + 255) / 256)), dim3(256), 0, s, a);
    if (a.nterms) hipLaunchKernelGGL(k_hsc_terms, dim3((a.nterms + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t scan_u32_to_u64(const uint32_t* in, uint64_t* out, uint64_t n, hipStream_t s)
{
    return scan_rec<uint32_t>(in, out, n, s);
}

}  // namespace exc
