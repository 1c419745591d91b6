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
//  * one wave owns one work unit = (query, 512 consecutive docIDs of its
//    shortest list, i.e. 4 segments of 128 or 2 of 256); the probe launches
//    one 64-thread workgroup per unit (units of the highest length-ratio class
//    first), each wave with its own LDS slice, synchronised with wave barriers
//    only.  Persistent launches claim chunks of units for the workgroup from
//    one device counter and take them one at a time through an LDS word.
//  * the unit's docIDs are decoded cooperatively into LDS and stay there as the
//    candidate set; each longer list (ascending length) filters it.
//  * per list, the heads of the blocks that can hold a candidate: for the
//    second list the prologue kernel's bounds, otherwise one gather of <= 64
//    heads around the window the list-level fit predicts (checked against the
//    heads, exact fallback to a 64-ary search);
//  * then either the wave decodes that run of blocks straight from registers
//    into a bitmap of a docID window in LDS and every candidate tests its bit
//    (low length ratios), or each row of 64 candidates finds its blocks among
//    the heads, stages the words of the positions each block's fit allows in
//    LDS and binary-searches them there (high ratios); ParaPFD, which has no
//    random access, decodes the run into the bitmap the same way, or a
//    candidate's half block by one row of 16 lanes (the block's median picks
//    the half);
//  * survivors are compacted with ballot + mbcnt (order preserving), so the
//    next list only probes live candidates and no 0/1 flag array exists;
//  * each unit writes its survivors into the stage slot of its shortest-list
//    segment; a block scan over unit counts + a scatter produce the contiguous
//    per-query output and cumulative counts.
#include "exc_format.h"
#include "exc_internal.h"
#include "exc_model.h"

namespace exc {
namespace {

constexpr int kWaves = 4;
constexpr int kProbeWaves = 1;  // waves (work units in flight) per probe workgroup
constexpr uint32_t kNone = 0xffffffffu;
constexpr uint32_t kClaim = 8;        // units per workgroup claim on the device counter (persistent launch)
// ParaPFD probes take the sparse path (one candidate's half block per row of
// 16 lanes) when the candidates average at most this many per distinct block
constexpr uint32_t kPfdSparse = 4;
constexpr uint32_t kBufWords = 768;  // per-wave LDS decode buffer (holds the LRC/SegLRC bitmap of 16,384 docIDs)
constexpr uint32_t kMaxWindows = 8;   // bitmap windows per cooperative ParaPFD probe
constexpr uint32_t kMaxWindowsSeg = 12;  // bitmap windows per cooperative LRC/SegLRC probe (16,384 docIDs each)
constexpr uint32_t kBitmapPiece = 768; // docIDs decoded per load round of the bitmap path
constexpr uint32_t kWinWords = 12;     // words of a candidate's fit window staged in LDS (64 x 12 <= kBufWords)
static_assert(64 * kWinWords <= kBufWords, "window staging fits the decode buffer");
constexpr int kV = (int)(kChunk / 64);  // candidates per lane

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

// A wave-uniform float copied into a vector register: a VALU instruction with
// two scalar operands needs one of them moved first (one scalar operand per
// instruction on gfx9), so a value used that way across a loop is moved once.
__device__ __forceinline__ float vgpr_copy(float x)
{
    float r;
    asm("v_mov_b32 %0, %1" : "=v"(r) : "s"(x));
    return r;
}

// The LRC/SegLRC window bitmap: kBmWords words at a 2 KiB-aligned LDS
// address inside the wave's buffer, docID offset x (< 32 * kBmWords) in word
// x % kBmWords, bit x / kBmWords.  Consecutive docIDs land in different words
// (and banks), so the atomic ORs of a dense run of values do not queue on one
// word as they do with word x / 32, bit x % 32; a power-of-two word count
// makes the position a mask and a shift (with 768 words it took a division
// by 768, five instructions per value), and the aligned base makes the word
// address one shift-or.  The ParaPFD bitmap keeps the plain layout: its values
// are gap sums, seldom adjacent.
constexpr uint32_t kBmWords = 512;
// The bitmap inside a wave's buffer (kBufWords words): the buffer's first
// 2 KiB boundary, or nullptr if the bitmap would run past the buffer's end
// (k_probe aligns its LDS to 2 KiB and every wave's buffer starts at a
// multiple of 1 KiB, so this does not happen).
__device__ __forceinline__ uint32_t* bitmap_of(uint32_t* buf)
{
    const uint32_t pad = (0u - (uint32_t)(uintptr_t)buf) & 2047u;
    return uni(pad) + 4u * kBmWords <= 4u * kBufWords ? buf + pad / 4u : nullptr;
}

// x >> 5 as one instruction the compiler cannot look through: it otherwise
// rewrites ((x >> 5) << 2) + base as ((x >> 3) & ~3) + base, three
// instructions where a shift and one v_lshl_add suffice.
__device__ __forceinline__ uint32_t word_of_bit(uint32_t x)
{
    uint32_t r;
    asm("v_lshrrev_b32 %0, 5, %1" : "=v"(r) : "v"(x));
    return r;
}

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

// Inclusive sum over the wave's lanes in DPP moves (VALU only, no LDS-unit
// permutes): row_shr 1, 2, 4, 8 within each row of 16 lanes (sources before
// the row start read 0), then rows 1 and 3 add lane 15 of the row before
// (row_bcast:15) and rows 2 and 3 add lane 31 (row_bcast:31).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint32_t dpp_add(uint32_t v)
{
    return v + (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xf, false);
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    v = dpp_add<0x111, 0xf>(v);  // row_shr:1
    v = dpp_add<0x112, 0xf>(v);  // row_shr:2
    v = dpp_add<0x114, 0xf>(v);  // row_shr:4
    v = dpp_add<0x118, 0xf>(v);  // row_shr:8
    v = dpp_add<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v = dpp_add<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
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

// wave_upper_bound of two values at once: both searches advance a level per
// load round, so their dependent loads overlap (one chain of rounds, not two).
__device__ void wave_upper_bound2(const uint32_t* __restrict__ a, uint32_t n, uint32_t va, uint32_t vb, uint32_t& ra,
                                  uint32_t& rb)
{
    const uint32_t ln = lane_id();
    uint32_t lo_a = 0, hi_a = n, lo_b = 0, hi_b = n;
    bool done_a = false, done_b = false;
    ra = rb = 0;
    while (!done_a || !done_b) {
        // a side whose range is <= 64 wide takes its last level this round
        const bool last_a = hi_a - lo_a <= 64, last_b = hi_b - lo_b <= 64;
        const uint32_t st_a = last_a ? 1u : (hi_a - lo_a + 63) >> 6, st_b = last_b ? 1u : (hi_b - lo_b + 63) >> 6;
        const uint32_t ia = lo_a + ln * st_a, ib = lo_b + ln * st_b;
        const bool pa = !done_a && ia < hi_a && a[ia] <= va;
        const bool pb = !done_b && ib < hi_b && a[ib] <= vb;
        const uint32_t ca = popc(__ballot(pa)), cb = popc(__ballot(pb));
        if (!done_a) {
            if (last_a || ca == 0) {
                ra = lo_a + (last_a ? ca : 0u);
                done_a = true;
            } else {
                const uint32_t nlo = lo_a + (ca - 1) * st_a + 1, nhi = min(lo_a + ca * st_a, hi_a);
                lo_a = uni(nlo);
                hi_a = uni(nhi);
            }
        }
        if (!done_b) {
            if (last_b || cb == 0) {
                rb = lo_b + (last_b ? cb : 0u);
                done_b = true;
            } else {
                const uint32_t nlo = lo_b + (cb - 1) * st_b + 1, nhi = min(lo_b + cb * st_b, hi_b);
                lo_b = uni(nlo);
                hi_b = uni(nhi);
            }
        }
    }
    ra = uni(ra);
    rb = uni(rb);
}

// Count of lanes of the sorted register array H (kNone padded) that are <= v,
// v per lane: 6 shuffles of binary lifting plus the 64th lane.
__device__ __forceinline__ uint32_t lane_count_le(uint32_t H, uint32_t h63, uint32_t v)
{
    uint32_t pos = 0;
#pragma unroll
    for (uint32_t st = 32; st >= 1; st >>= 1) {
        const uint32_t h = __shfl(H, (int)(pos + st - 1));
        if (h <= v) pos += st;
    }
    if (pos == 63 && h63 <= v) pos = 64;  // the lifting steps sum to 63
    return pos;
}

__device__ __forceinline__ uint32_t seg_len(const TermRec& R, uint32_t s, uint32_t seg)
{
    return (s + 1 == R.nseg) ? R.n - s * seg : seg;
}

// ------------------------------------------------------------ block decode

// d-gap codecs: blocks decode by exception patch + prefix sum from segHead
constexpr bool gap_codec(int codec) { return codec == EXC_PARAPFD || codec == EXC_NEWPFD; }

// Raw lb-bit fields of one block for this lane: strided positions
// (j = lane + 64 t) for the random-access codecs, consecutive (j = lane*V + t)
// for the d-gap codecs whose prefix sum runs along a lane first.
// Every lane loads every field position j < SEG, inside the block or not (an
// array's last block reads into its kPadWords of zero padding), so that all
// loads of a group are in flight before the first wait and no lane spends
// instructions clamping its position.  ZERO: fields past the block's length
// read as 0 (callers that test j < len themselves skip it).
template <int CODEC, int SEG, bool ZERO = true>
__device__ __forceinline__ void load_fields(const uint32_t* __restrict__ data, uint32_t lb, uint32_t len,
                                            uint32_t (&f)[SEG / 64])
{
    static_assert(SEG <= (int)kMaxSeg, "field reads past the block stay inside the padding");
    constexpr int V = SEG / 64;
    const uint32_t ln = lane_id();
#pragma unroll
    for (int t = 0; t < V; ++t) {
        const uint32_t j = gap_codec(CODEC) ? ln * V + t : ln + 64 * t;
        const uint32_t x = bits_at(data, lb, j);
        f[t] = (!ZERO || j < len) ? x : 0u;
    }
}

// The 16 Simple16 layouts (s16_layout), for the lane-shuffle lookup
__constant__ uint32_t kS16Layouts[16] = {s16_layout(0),  s16_layout(1),  s16_layout(2),  s16_layout(3),
                                         s16_layout(4),  s16_layout(5),  s16_layout(6),  s16_layout(7),
                                         s16_layout(8),  s16_layout(9),  s16_layout(10), s16_layout(11),
                                         s16_layout(12), s16_layout(13), s16_layout(14), s16_layout(15)};

// The exceptions of a NewPFor block patched into dst[0..len) (low parts
// already there).  The reference decodes the Simple16 stream and patches in
// thread 0 alone (GC/NewPFD/NewPFD_Decompression/NewPFD_Decompression_kernel.cu:420-447);
// here each lane takes one word of the stream, a wave scan of the words'
// field counts gives every word its first field, each lane then extracts the
// fields l, l + 64, ... of xs[0 .. 2 en) from the words that hold them, a
// scan of the position deltas gives every exception its position, and the
// patch is lane-parallel.
// Stream words are read 64 at a time: past the stream's end a round reads at
// most 63 words, inside the list's successors or the array's zero padding.
// w_first: this lane's word of the first round, loaded with the block's fields.
// s16t: this lane's entry of the Simple16 layout table (selector lane % 16,
// kS16Layouts); a word's layout is then one lane shuffle, where a per-lane
// switch on the selector compiles to a divergent decision tree of some 60
// branch-masked instructions.
__device__ __forceinline__ void npfd_patch(const uint32_t* __restrict__ sp, uint32_t en, uint32_t lb, uint32_t len,
                                           uint32_t w_first, uint32_t* xs, uint32_t* dst, uint32_t s16t)
{
    const uint32_t ln = lane_id(), need = 2 * en;
    for (uint32_t w0 = 0, got = 0; got < need; w0 += 64) {
        const uint32_t word = w0 ? sp[w0 + ln] : w_first;
        const uint32_t c = s16_count_l((uint32_t)__shfl((int)s16t, (int)(word >> 28)));
        const uint32_t incl = wave_incl_scan(c);
        const uint32_t ex = incl - c;  // the round's first field in this lane's word
        const uint32_t tot = uni(__builtin_amdgcn_readlane(incl, 63));
        const uint32_t nf = min(tot, need - got);
        // field-parallel unpack: lane l takes fields l, l + 64, ... of the
        // round; its word is the last lane whose first field is <= f (binary
        // lifting over the lanes, ex strictly increases since every word holds
        // a field), so no lane loops over a long word while the others idle
        for (uint32_t f0 = 0; f0 < nf; f0 += 64) {
            const uint32_t f = f0 + ln;
            uint32_t o = 0;
            for (uint32_t step = 32; step; step >>= 1)
                if ((uint32_t)__shfl((int)ex, (int)(o + step)) <= f) o += step;
            const uint32_t wo = (uint32_t)__shfl((int)word, (int)o);
            const uint32_t k = f - (uint32_t)__shfl((int)ex, (int)o);
            const uint32_t v = s16_field_l(wo, k, (uint32_t)__shfl((int)s16t, (int)(wo >> 28)));
            if (f < nf) xs[got + f] = v;
        }
        got += tot;
    }
    wave_sync();
    // position i = all[0] + sum over m = 1..i of (all[m] + 1)  (GI/NewPFor.cpp:119-124)
    for (uint32_t e0 = 0, carry = 0; e0 < en; e0 += 64) {
        const uint32_t e = e0 + ln;
        const uint32_t d = e < en ? xs[e] + (e ? 1u : 0u) : 0u;
        const uint32_t p = wave_incl_scan(d) + carry;
        if (e < en && p < len)
            __hip_atomic_fetch_add(dst + p, xs[en + e] << lb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        carry = uni(__builtin_amdgcn_readlane(p, 63));
    }
}

// The first exception words of a d-gap block, loaded with its fields so that
// the patch needs no second round trip: ParaPFD exception e = lane's position
// and high part, NewPFD the lane's word of the first Simple16 round.
struct ExcPrefetch {
    uint32_t a = 0, b = 0;
};

template <int CODEC>
__device__ __forceinline__ ExcPrefetch prefetch_exceptions(const uint32_t* __restrict__ data, uint32_t desc, uint32_t len)
{
    ExcPrefetch x;
    const uint32_t en = desc_en(desc);
    if (!gap_codec(CODEC) || !en) return x;
    const uint32_t* xp = data + words_for(len, desc_lb(desc));
    const uint32_t ln = lane_id();
    if constexpr (CODEC == EXC_NEWPFD) {
        x.a = xp[ln];
    } else {
        const uint32_t eb = desc_eb(desc), e = min(ln, en - 1u);
        x.a = bits_at(xp, eb, e);
        x.b = bits_at(xp + words_for(en, eb), desc_hb(desc), e);
    }
    return x;
}

// The exceptions of a ParaPFor block whose fields are staged in dst[0..len)
// ORed in place, lane-parallel (GI/ParaPFor.cpp:107-126: field[pos_e] |=
// high_e << lb; positions are distinct, so no two lanes write one word).
// PREFETCH: exception e < 64's position and high part arrived in xp with the
// fields.  Positions at or past len (none in a valid block) are ignored.
// Every caller's dst holds at least 64 words (a whole block's stage).
template <bool PREFETCH>
__device__ __forceinline__ void pfd_patch_lds(const uint32_t* __restrict__ data, uint32_t desc, uint32_t len,
                                              ExcPrefetch xp, uint32_t* dst)
{
    const uint32_t lb = desc_lb(desc), en = desc_en(desc), eb = desc_eb(desc), hb = desc_hb(desc);
    const uint32_t* pp = data + words_for(len, lb);
    const uint32_t* hp = pp + words_for(en, eb);
    const uint32_t ln = lane_id();
    if (!en) return;
    // exceptions 0..63 (prefetched with the fields, or read here): one round
    // with no lane branching, lanes without an exception ORing 0 into their
    // own word
    if constexpr (!PREFETCH) {
        const uint32_t e = min(ln, en - 1u);
        xp.a = bits_at(pp, eb, e);
        xp.b = bits_at(hp, hb, e);
    }
    const bool act = ln < en && xp.a < len;
    __hip_atomic_fetch_or(dst + (act ? xp.a : ln), act ? xp.b << lb : 0u, __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_WAVEFRONT);
    if (en <= 64) return;
    for (uint32_t e = 64 + ln; e < en; e += 64) {
        const uint32_t p = bits_at(pp, eb, e), h = bits_at(hp, hb, e);
        // an LDS atomic OR (no return): no read round trip before the write
        if (p < len) __hip_atomic_fetch_or(dst + p, h << lb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    }
}

// Finish decoding block s (fields already loaded) into dst[0..len); xs is
// the NewPFor exception scratch (2 SEG words), xp the prefetched exceptions.
template <int CODEC, int SEG, bool PREFETCH = false>
__device__ __forceinline__ void finish_block(const TermRec& R, uint32_t s, const SegRec& rec, uint32_t len,
                                             const uint32_t* __restrict__ data, uint32_t (&f)[SEG / 64],
                                             uint32_t* dst, uint32_t* xs, ExcPrefetch xp = ExcPrefetch(),
                                             uint32_t s16t = 0)
{
    constexpr int V = SEG / 64;
    const uint32_t ln = lane_id();
    const uint32_t lb = desc_lb(rec.desc);
    if constexpr (CODEC == EXC_LRC) {
        // GC/LRC/LRC_Decompression/LRC_Decompression_kernel.cu:206-256
#pragma unroll
        for (int t = 0; t < V; ++t) {
            const uint32_t j = ln + 64 * t;
            if (j < len) dst[j] = f[t] + lrc_pred(s * SEG + j, R.slope, R.icpt);
        }
    } else if constexpr (CODEC == EXC_SEGLRC) {
        // GC/SegLRC/SegLRC_Decompression/SegLRC_Decompression_kernel.cu:200-256
        // (the intercept moved to a vector register once: seg_pred's fma then
        // reads one scalar operand)
        const float icv = vgpr_copy(rec.icpt);
#pragma unroll
        for (int t = 0; t < V; ++t) {
            const uint32_t j = ln + 64 * t;
            if (j < len) dst[j] = f[t] + seg_pred(j, rec.slope, icv, rec.base);
        }
    } else {
        // ParaPFD / NewPFD: patch exceptions, segHead, inclusive scan
        // (GC/ParaPFD/ParaPFD_Decompression/ParaPFD_Decompression_kernel.cu:198-264,
        // GC/NewPFD/NewPFD_Decompression/NewPFD_Decompression_kernel.cu:405-470).
        const uint32_t en = desc_en(rec.desc);
        if (en) {
#pragma unroll
            for (int t = 0; t < V; ++t) {
                const uint32_t j = ln * V + t;
                if (j < len) dst[j] = f[t];
            }
            wave_sync();
            if constexpr (CODEC == EXC_NEWPFD) {
                const uint32_t* sp = data + words_for(len, lb);
                npfd_patch(sp, en, lb, len, PREFETCH ? xp.a : sp[ln], xs, dst, s16t);
            } else {
                pfd_patch_lds<PREFETCH>(data, rec.desc, len, xp, dst);
            }
            wave_sync();
#pragma unroll
            for (int t = 0; t < V; ++t) {
                const uint32_t j = ln * V + t;
                if (j < len) f[t] = dst[j];
            }
        }
        if (ln == 0) f[0] = rec.head;
        uint32_t acc = 0;
#pragma unroll
        for (int t = 0; t < V; ++t) { acc += f[t]; f[t] = acc; }
        const uint32_t excl = wave_incl_scan(acc) - acc;
        wave_sync();
#pragma unroll
        for (int t = 0; t < V; ++t) {
            const uint32_t j = ln * V + t;
            if (j < len) dst[j] = f[t] + excl;
        }
    }
}

// Decode segments listed by sid(g), g < r <= NB, of list R into dst + g*SEG
// (sid is called with a lane-dependent g); NewPFD needs 2 SEG words of
// scratch after them, at dst + NB*SEG.  LANE_RECS: a group's segment
// records come in one lane-parallel load and are spread by readlane (fewer
// registers: the whole-list decoder); otherwise one uniform load per block
// (what the probe kernel's register budget prefers).
// Blocks go in groups (all of them for the random-access codecs, pairs for
// ParaPFD whose scans hold more registers); every load of a group is issued
// before any of its values is consumed.  PREFETCH: the d-gap codecs' first
// exception words are loaded with the fields (two more registers per block:
// the whole-list decoder takes them, the probe kernel's ParaPFD path not).
template <int CODEC, int SEG, int NB, bool LANE_RECS = false, bool PREFETCH = false, class SegId>
__device__ void decode_blocks(const DevIndex& ix, const TermRec& R, uint32_t r, SegId sid, uint32_t* dst)
{
    constexpr int V = SEG / 64;
    constexpr int GG = PREFETCH ? 4 : 2;  // d-gap blocks in flight: the decoder's register budget allows more
    constexpr int G = gap_codec(CODEC) ? (NB < GG ? NB : GG) : NB;
    // LANE_RECS: lane l < NB holds block min(l, r - 1)'s record, all in one load
    uint32_t sl_all = 0;
    SegRec gr_all;
    if constexpr (LANE_RECS) {
        sl_all = sid((int)min(min(lane_id(), (uint32_t)(NB - 1)), r - 1));
        gr_all = ix.rec[R.segnum + sl_all];
    }
    // NewPFD: this lane's Simple16 layout-table entry, once per call
    const uint32_t s16t = CODEC == EXC_NEWPFD ? kS16Layouts[lane_id() & 15u] : 0u;
    for (uint32_t g0 = 0; g0 < r; g0 += G) {
        SegRec rec[G];
        uint32_t len[G], sg[G];
        const uint32_t* data[G];
        uint32_t f[G][V];
        // past r: a copy of the last block, loaded but not used
        if constexpr (LANE_RECS) {
            const uint32_t sl = sl_all;
            const SegRec gr = gr_all;
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const uint32_t src = min(g0 + (uint32_t)g, r - 1);
                sg[g] = uni(__builtin_amdgcn_readlane(sl, src));
                rec[g].off = uni(__builtin_amdgcn_readlane(gr.off, src));
                rec[g].desc = uni(__builtin_amdgcn_readlane(gr.desc, src));
                rec[g].slope = __uint_as_float(uni(__builtin_amdgcn_readlane(__float_as_uint(gr.slope), src)));
                rec[g].icpt = __uint_as_float(uni(__builtin_amdgcn_readlane(__float_as_uint(gr.icpt), src)));
                rec[g].rl = rec[g].rr = 0;
                rec[g].base = (int32_t)uni(__builtin_amdgcn_readlane((uint32_t)gr.base, src));
                rec[g].head = uni(__builtin_amdgcn_readlane(gr.head, src));
            }
        } else {
#pragma unroll
            for (int g = 0; g < G; ++g) {
                sg[g] = uni(sid((int)min(g0 + g, r - 1)));
                rec[g] = ix.rec[R.segnum + sg[g]];
            }
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
            len[g] = seg_len(R, sg[g], SEG);
            data[g] = ix.words + R.woff + rec[g].off + 1;
        }
        ExcPrefetch xp[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            load_fields<CODEC, SEG>(data[g], desc_lb(rec[g].desc), len[g], f[g]);
            if constexpr (PREFETCH) xp[g] = prefetch_exceptions<CODEC>(data[g], rec[g].desc, len[g]);
        }
#pragma unroll
        for (int g = 0; g < G; ++g)
            if (g0 + g < r)
                finish_block<CODEC, SEG, PREFETCH>(R, sg[g], rec[g], len[g], data[g], f[g], dst + (g0 + g) * SEG,
                                                   dst + NB * SEG, xp[g], s16t);
    }
}

// The blocks s0 .. s0+r-1 (r <= NB) of a ParaPFD list decoded into dst
// (block g at dst + g * SEG), phase by phase over all of them: every block's
// fields and first 64 exceptions are requested in one load round, all are
// staged in dst, patched, read back, scanned and written, so their LDS round
// trips overlap instead of forming one chain per block (the whole-list
// decoder showed the same ordering 28 % faster).  Exceptions past the first
// 64 of a block are read in pfd_patch_lds's loop.
template <int SEG, int NB>
__device__ void decode_pfd_phased(const DevIndex& ix, const TermRec& R, uint32_t s0, uint32_t r, uint32_t* dst)
{
    constexpr int V = SEG / 64;
    const uint32_t ln = lane_id();
    const SegRec gr = ix.rec[R.segnum + s0 + min(ln, r - 1)];
    uint32_t desc[NB], head[NB], len[NB], f[NB][V];
    const uint32_t* data[NB];
    ExcPrefetch xp[NB];
#pragma unroll
    for (int g = 0; g < NB; ++g) {
        const int src = (int)min((uint32_t)g, r - 1);  // past r: a copy of the last block, not used
        desc[g] = uni(__builtin_amdgcn_readlane(gr.desc, src));
        head[g] = uni(__builtin_amdgcn_readlane(gr.head, src));
        len[g] = seg_len(R, s0 + (uint32_t)src, SEG);
        data[g] = ix.words + R.woff + uni(__builtin_amdgcn_readlane(gr.off, src)) + 1;
        load_fields<EXC_PARAPFD, SEG, false>(data[g], desc_lb(desc[g]), len[g], f[g]);
        xp[g] = prefetch_exceptions<EXC_PARAPFD>(data[g], desc[g], len[g]);
    }
    bool exc = false;
#pragma unroll
    for (int g = 0; g < NB; ++g) exc = exc || ((uint32_t)g < r && desc_en(desc[g]));
    if (exc) {
#pragma unroll
        for (int g = 0; g < NB; ++g)
#pragma unroll
            for (int t = 0; t < V; ++t) dst[g * SEG + ln * V + t] = f[g][t];
        wave_sync();
#pragma unroll
        for (int g = 0; g < NB; ++g)
            if ((uint32_t)g < r && desc_en(desc[g])) pfd_patch_lds<true>(data[g], desc[g], len[g], xp[g], dst + g * SEG);
        wave_sync();
#pragma unroll
        for (int g = 0; g < NB; ++g)
#pragma unroll
            for (int t = 0; t < V; ++t) f[g][t] = dst[g * SEG + ln * V + t];
    }
#pragma unroll
    for (int g = 0; g < NB; ++g) {
        if (ln == 0) f[g][0] = head[g];
        uint32_t acc = 0;
#pragma unroll
        for (int t = 0; t < V; ++t) { acc += f[g][t]; f[g][t] = acc; }
        const uint32_t excl = wave_incl_scan(acc) - acc;
#pragma unroll
        for (int t = 0; t < V; ++t) f[g][t] += excl;
    }
    wave_sync();
#pragma unroll
    for (int g = 0; g < NB; ++g)
#pragma unroll
        for (int t = 0; t < V; ++t) {
            const uint32_t j = ln * V + t;
            if ((uint32_t)g < r && j < len[g]) dst[g * SEG + j] = f[g][t];
        }
}

// ------------------------------------------------------------ random access

__device__ __forceinline__ uint32_t seg_val(const uint32_t* data, uint32_t lb, uint32_t j, const SegRec& r)
{
    return bits_at(data, lb, j) + seg_pred(j, r.slope, r.icpt, r.base);
}

__device__ __forceinline__ bool seg_bsearch(const uint32_t* data, uint32_t lb, const SegRec& r, int lo, int hi,
                                            uint32_t v)
{
    while (lo <= hi) {
        const int m = (lo + hi) >> 1;
        const uint32_t q = seg_val(data, lb, (uint32_t)m, r);
        if (q == v) return true;
        if (q < v) lo = m + 1; else hi = m - 1;
    }
    return false;
}

// Exact search of v in a random-access SegLRC / HS block: start at the
// position the block's fit predicts, walk towards v for at most the fit's
// recorded deviation (rl / rr), and binary-search the rest of the block if v
// may lie beyond the walk.  The reference searches only the fit window
// (HS_SegLRC_Intersection_kernel.cu:361-369, SegLRC_Intersection_kernel.cu:366-371),
// which drops results; the fallback keeps the answer exact.
__device__ __forceinline__ bool seg_find(const uint32_t* data, uint32_t len, const SegRec& r, uint32_t v, bool window)
{
    const uint32_t lb = desc_lb(r.desc);
    if (!window) return seg_bsearch(data, lb, r, 0, (int)len - 1, v);
    const float fp = ((float)v - r.icpt) * __builtin_amdgcn_rcpf(r.slope) + 0.5f;
    int p = (int)fminf(fmaxf(fp, 0.0f), (float)(len - 1));
    uint32_t q = seg_val(data, lb, (uint32_t)p, r);
    if (q == v) return true;
    if (q < v) {
        const int lim = min((int)len - 1, p + (int)min(r.rr, 64u) + 1);
        while (p < lim) {
            ++p;
            q = seg_val(data, lb, (uint32_t)p, r);
            if (q >= v) return q == v;
        }
        return p < (int)len - 1 && seg_bsearch(data, lb, r, p + 1, (int)len - 1, v);
    }
    const int lim = max(0, p - (int)min(r.rl, 64u) - 1);
    while (p > lim) {
        --p;
        q = seg_val(data, lb, (uint32_t)p, r);
        if (q <= v) return q == v;
    }
    return p > 0 && seg_bsearch(data, lb, r, 0, p - 1, v);
}

// Value j of block s of a random-access list from its residual f
// (the expressions of finish_block).
template <int CODEC, int SEG>
__device__ __forceinline__ uint32_t block_value(const TermRec& R, uint32_t s, const SegRec& rec, uint32_t j, uint32_t f)
{
    if constexpr (CODEC == EXC_LRC)
        return f + lrc_pred(s * SEG + j, R.slope, R.icpt);
    else
        return f + seg_pred(j, rec.slope, rec.icpt, rec.base);
}

// The blocks s0 .. s0+r-1 (r <= NB) of an LRC / SegLRC list decoded into dst
// (block g at dst + g * SEG), dst holding NB * SEG words: every lane stores
// every position of every block, past a block's length and past r too (a
// copy of the last block), so no store waits behind a per-lane branch; the
// words past the run are never read as docIDs (the probe's candidate count
// stops before them).  Records in scalar registers, one load each, all
// requested before the first field.
template <int CODEC, int SEG, int NB>
__device__ void decode_ra_padded(const DevIndex& ix, const TermRec& R, uint32_t s0, uint32_t r, uint32_t* dst)
{
    constexpr int V = SEG / 64;
    const uint32_t ln = lane_id();
    SegRec rec[NB];
    uint32_t f[NB][V];
#pragma unroll
    for (int g = 0; g < NB; ++g) rec[g] = ix.rec[R.segnum + s0 + min((uint32_t)g, r - 1)];
#pragma unroll
    for (int g = 0; g < NB; ++g)
        load_fields<CODEC, SEG, false>(ix.words + R.woff + rec[g].off + 1, desc_lb(rec[g].desc), SEG, f[g]);
#pragma unroll
    for (int g = 0; g < NB; ++g) {
        const uint32_t s = s0 + min((uint32_t)g, r - 1);
        const float icv = vgpr_copy(rec[g].icpt);
#pragma unroll
        for (int t = 0; t < V; ++t) {
            const uint32_t j = ln + 64 * t;
            uint32_t x;
            if constexpr (CODEC == EXC_SEGLRC)
                x = f[g][t] + seg_pred(j, rec[g].slope, icv, rec[g].base);
            else
                x = block_value<CODEC, SEG>(R, s, rec[g], j, f[g][t]);
            dst[g * SEG + j] = x;
        }
    }
}

// Value j of a SegLRC block whose packed words from w0 on are staged in LDS,
// word k at col[64 k] (one column per lane).
__device__ __forceinline__ uint32_t win_value(const uint32_t* col, uint32_t lb, uint32_t w0, uint32_t j,
                                              const SegRec& r)
{
    const uint32_t bit = j * lb - w0 * 32u, k = bit >> 5;
    const uint64_t cw = (uint64_t)col[64u * k] | ((uint64_t)col[64u * (k + 1)] << 32);
    return ((uint32_t)(cw >> (bit & 31)) & ((1u << lb) - 1u)) + seg_pred(j, r.slope, r.icpt, r.base);
}

// Exact search of v in a random-access SegLRC / HS block, one candidate per
// lane, every lane taking part: the words of the positions the block fit
// allows ([p - rl - 1, p + rr + 1], up to kWinWords words) arrive in one load
// round, are staged in buf (word k of lane l at buf[64k + l]) and searched
// there.  A v outside the window's values, a window too wide or the binary
// search mode fall back to the search in HBM.  Lanes with act false return
// false (their data / r must still be a valid block).
__device__ bool win_search(uint32_t* buf, const uint32_t* data, uint32_t len, const SegRec& r, uint32_t v, bool act,
                           uint32_t search)
{
    const uint32_t ln = lane_id(), lb = desc_lb(r.desc);
    const float fp = ((float)v - r.icpt) * __builtin_amdgcn_rcpf(r.slope) + 0.5f;
    const int last = (int)len - 1;
    const int pp = (int)fminf(fmaxf(fp, 0.0f), (float)max(last, 0));
    const int wl = max(0, pp - (int)min(r.rl, 64u) - 1), wh = max(wl, min(last, pp + (int)min(r.rr, 64u) + 1));
    const uint32_t w0 = ((uint32_t)wl * lb) >> 5;
    const bool ok = search != SEARCH_BINARY && lb > 0 && len > 0 && ((((uint32_t)wh * lb) >> 5) - w0) + 2 <= kWinWords;
    wave_sync();
#pragma unroll
    for (int k = 0; k < (int)kWinWords; ++k) buf[64 * k + ln] = data[w0 + k];
    wave_sync();
    if (!act || len == 0) return false;
    if (!ok) return seg_find(data, len, r, v, search != SEARCH_BINARY);
    const uint32_t vlo = win_value(buf + ln, lb, w0, (uint32_t)wl, r), vhi = win_value(buf + ln, lb, w0, (uint32_t)wh, r);
    // v inside the window's values: the last position with a value <= v, in
    // straight-line steps of halving powers of two (the wave's widest window
    // sets the first), so that no lane branches out of a search loop
    const bool inwin = v >= vlo && v <= vhi;
    const uint32_t n = inwin ? (uint32_t)(wh - wl) : 0u;
    uint32_t step = 128;  // windows span at most 2 x 64 + 3 positions
    while (step > 1 && !__ballot(n >= step)) step >>= 1;
    int pos = wl;
    uint32_t vp = vlo;
    for (; step; step >>= 1) {
        const int m = pos + (int)step;
        const uint32_t q = win_value(buf + ln, lb, w0, (uint32_t)min(m, wh), r);
        const bool adv = m <= wh && q <= v;
        pos = adv ? m : pos;
        vp = adv ? q : vp;
    }
    if (inwin) return vp == v;
    if (v < vlo) return wl > 0 && seg_bsearch(data, lb, r, 0, wl - 1, v);
    return wh < last && seg_bsearch(data, lb, r, wh + 1, last, v);
}

// Branchless search of v in the sorted LDS run buf[0..n) (n uniform).
__device__ __forceinline__ bool lds_contains(const uint32_t* buf, uint32_t n, uint32_t v)
{
    uint32_t pos = 0;
    for (uint32_t step = 1u << (31 - __builtin_clz(n | 1u)); step; step >>= 1)
        if (pos + step <= n && buf[pos + step - 1] < v) pos += step;
    return pos < n && buf[pos] == v;
}

// LRC random access: binary search of the segment (no per-segment fit)
// (GC/LRC/LRC_Intersection/LRC_Intersection_kernel.cu:333-355).
__device__ __forceinline__ bool lrc_bsearch(const uint32_t* data, uint32_t lb, uint32_t base_index, uint32_t len,
                                            const TermRec& R, uint32_t v)
{
    int lo = 0, hi = (int)len - 1;
    while (lo <= hi) {
        const int m = (lo + hi) >> 1;
        const uint32_t q = bits_at(data, lb, (uint32_t)m) + lrc_pred(base_index + (uint32_t)m, R.slope, R.icpt);
        if (q == v) return true;
        if (q < v) lo = m + 1; else hi = m - 1;
    }
    return false;
}

// ------------------------------------------------------------ per-list probes

// Window [wa, wb] for "number of segment heads <= v" from the list-level fit.
template <int CODEC, int SEG>
__device__ __forceinline__ void fit_window(const TermRec& R, uint32_t v, int64_t& wa, int64_t& wb)
{
    const float icpt = (CODEC == EXC_LRC) ? (float)R.icpt : __int_as_float(R.icpt);
    float p = ((float)v - icpt) * __builtin_amdgcn_rcpf(R.slope);
    p = fminf(fmaxf(p, -1.0e9f), 4.0e9f);
    const float plo = p - R.range_l - 2.0f, phi = p + R.range_r + 2.0f;
    wa = plo <= 0.0f ? 0 : (int64_t)(plo / SEG);
    wb = phi <= 0.0f ? 1 : (int64_t)(phi / SEG) + 2;
}

// mlp_pred for vmin and vmax at once: lane j (mod 32) evaluates hidden unit j,
// lanes 0-31 for vmin and 32-63 for vmax; the output sums then run in unit
// order from the lanes' values, the same operations as mlp_pred, so the
// result is bit-identical while each lane holds three parameters.
__device__ __forceinline__ void mlp_pred_pair(const float* w, float x0, float xs, uint32_t n, uint32_t vmin,
                                              uint32_t vmax, float& plo, float& phi)
{
    const uint32_t ln = lane_id(), j = ln & (kMlpH - 1);
    const float x = ((float)(ln < 32 ? vmin : vmax) - x0) * xs;
    const float h = fmaxf(fmaf(w[j], x, w[kMlpH + j]), 0.0f);
    const float w2 = w[2 * kMlpH + j];
    const int hb = __float_as_int(h), wb = __float_as_int(w2);
    float a0 = w[3 * kMlpH], a1 = a0;
    for (int k = 0; k < kMlpH; ++k) {
        a0 = fmaf(__int_as_float(__builtin_amdgcn_readlane(wb, k)), __int_as_float(__builtin_amdgcn_readlane(hb, k)), a0);
        a1 = fmaf(__int_as_float(__builtin_amdgcn_readlane(wb, 32 + k)),
                  __int_as_float(__builtin_amdgcn_readlane(hb, 32 + k)), a1);
    }
    plo = a0 * (float)n;
    phi = a1 * (float)n;
}

// Window [wa, wb] for "number of segment heads <= v" over [vmin, vmax] from
// the list's learned model: prediction +- its recorded max error.
// (The model's scale, offset and error travel in the list's TermRec, so only
// the parameters are read here.)
template <int SEG>
__device__ __forceinline__ void model_window(const DevIndex& ix, const TermRec& R, uint32_t vmin, uint32_t vmax,
                                             int64_t& wa, int64_t& wb)
{
    const float* prm = ix.mparams + (uint64_t)R.model * model_stride(R.mkind);
    float plo, phi;
    if (R.mkind == kModelPla) {
        plo = pla_pred(prm, R.mx0, R.mxs, vmin);
        phi = pla_pred(prm, R.mx0, R.mxs, vmax);
    } else if (R.mkind == kModelPlaQ) {
        plo = plaq_pred(prm, vmin);
        phi = plaq_pred(prm, vmax);
    } else {
        mlp_pred_pair(prm, R.mx0, R.mxs, R.n, vmin, vmax, plo, phi);
    }
    plo = fminf(fmaxf(plo - R.merr - 2.0f, -1.0e9f), 4.0e9f);
    phi = fminf(fmaxf(phi + R.merr + 2.0f, -1.0e9f), 4.0e9f);
    wa = plo <= 0.0f ? 0 : (int64_t)(plo / SEG);
    wb = phi <= 0.0f ? 1 : (int64_t)(phi / SEG) + 2;
}

// Per-wave LDS: the unit's live candidates, then a buffer for decoded
// segments or a docID-window bitmap.  The ParaPFD distinct-segment path keeps
// its list of segments in the buffer's last kChunk words and decodes into the
// first kPfdDecWords, so every compressed codec fits 1,280 words per wave
// (eight waves per SIMD as far as LDS goes).
constexpr uint32_t kPfdDecWords = kBufWords - kChunk;
struct LdsWave {
    uint32_t* cand;   // [kChunk]    live candidates (sorted)
    uint32_t* buf;    // [kBufWords] decoded segments / window bitmap
    uint32_t* dlist;  // [kChunk]    = buf + kPfdDecWords: distinct segments (ParaPFD long-span path only)
};

template <int CODEC, int SEG>
__device__ uint32_t probe_seg(const ProbeArgs& A, const TermRec& R, const LdsWave& W, uint32_t nlive, uint32_t knobs,
                              uint32_t pre_first, uint32_t pre_count, uint32_t pre_heads)
{
    const uint32_t search = knob_search(knobs);
    constexpr int V = kV;
    constexpr int NB = (int)(kBufWords / SEG);
    const uint32_t ln = lane_id();
    const uint32_t* head = A.ix.seg_head + R.segnum;
    uint32_t* cand = W.cand;
    const uint32_t vmin = uni(cand[0]), vmax = uni(cand[nlive - 1]);
    const int64_t nseg = R.nseg;

    // ---- heads around the candidates: H = head[hb0 + lane] (index -1 reads as 0)
    int64_t hb0 = 0;
    uint32_t H = kNone;
    bool small = false;
    int64_t rng = 0;
    if (pre_count != kUnset) {
        // blocks found by the prologue from the slice's bounding heads (a
        // superset of those the live candidates need)
        if (pre_count == 0) return 0;
        hb0 = pre_first;
        rng = pre_count;
        small = rng <= 64;
        H = pre_heads;  // head[hb0 + lane] for lanes below rng (small), loaded by run_unit
    } else if (search != SEARCH_BINARY && nseg > 64) {
        int64_t wa, t0, t1, wb;
        if (search == SEARCH_LEARNED && R.model != kNoModel) {
            model_window<SEG>(A.ix, R, vmin, vmax, wa, wb);
        } else {
            fit_window<CODEC, SEG>(R, vmin, wa, t0);
            fit_window<CODEC, SEG>(R, vmax, t1, wb);
        }
        wa = min(wa, nseg);
        wb = min(max(wb, wa), nseg);
        if (wb - wa + 2 <= 64) {
            hb0 = wa - 1;
            // lanes below lim read head[hb0 + lane] (lane 0 of hb0 = -1 reads
            // as 0): 32-bit lane arithmetic against uniform bounds
            const uint32_t lim = uni((uint32_t)min(nseg - hb0, wb - wa + 2));
            H = (hb0 < 0 && ln == 0) ? 0u : (ln < lim ? (head + hb0)[ln] : kNone);
            const uint32_t h0 = uni(__builtin_amdgcn_readlane(H, 0));
            const uint32_t hend = uni(__builtin_amdgcn_readlane(H, (int)(wb - wa + 1)));
            small = h0 <= vmin && hend > vmax;
        }
    }
    if (!small && pre_count == kUnset) {
        uint32_t ub_lo, ub_hi;
        wave_upper_bound2(head, (uint32_t)nseg, vmin, vmax, ub_lo, ub_hi);
        if (ub_hi == 0) return 0;  // every candidate precedes the list
        hb0 = ub_lo ? ub_lo - 1 : 0;
        rng = (int64_t)ub_hi - hb0;
        small = rng <= 64;
        H = (small && ln < (uint32_t)rng) ? (head + hb0)[ln] : kNone;
    }
    // segments spanned by the candidate set
    int32_t s_first, s_last;
    if (small) {
        s_first = (int32_t)(hb0 + (int64_t)popc(__ballot(H <= vmin)) - 1);
        s_last = (int32_t)(hb0 + (int64_t)popc(__ballot(H <= vmax)) - 1);
    } else {
        s_first = (int32_t)hb0;
        s_last = (int32_t)(hb0 + rng - 1);
    }
    s_first = max(s_first, 0);
    if (s_last < 0) return 0;
    const uint32_t r = uni((uint32_t)(s_last - s_first + 1));
    const uint32_t sf = uni((uint32_t)s_first);

    // candidates are re-read from LDS where needed instead of being held in
    // registers through the bitmap path (its register peak)
    // (the read is unconditional -- every slot below kChunk is a valid LDS
    // word -- and the result selected, so no lane branches round it)
    const auto cand_at = [&](int i) {
        const uint32_t e = ln + 64 * i;
        const uint32_t x = cand[e];
        return e < nlive ? x : kNone;
    };
    bool found[V];
#pragma unroll
    for (int i = 0; i < V; ++i) found[i] = false;

    const uint32_t cmax = knob_coop(knobs);
    constexpr uint32_t kWin = kBmWords * 32u;   // docIDs one LRC/SegLRC window bitmap covers
    constexpr uint32_t kWinPfd = (kBufWords - SEG) * 32u;  // ParaPFD: the buffer's tail stages a block
    constexpr int NBB = kBitmapPiece / SEG;       // blocks per load round of the bitmap path
    uint32_t* const bm = bitmap_of(W.buf);
    if (CODEC != EXC_PARAPFD && small && r <= cmax && ((vmax - vmin) / kWin) < kMaxWindowsSeg && bm != nullptr) {
        // ---- LRC/SegLRC: the run sf .. sf+r-1 decoded straight from
        // registers into a bitmap of a docID window in LDS; every candidate of
        // the window then tests its bit.  Windows of kWin docIDs cover
        // [vmin, vmax]; each decodes only the blocks that overlap it (the run's
        // heads tell which) and windows without candidates are skipped.
        // Independent LDS operations only, no per-candidate search chains.
        // Per block: the record in scalar registers (one scalar load), each
        // lane's bit offset ln * lb once (the second field of a lane sits
        // 64 * lb bits = 2 lb words further on, the same bit offset in its
        // word), then per value a funnel shift, a mask, the fit and the
        // insertion; only a list's last block tests j < len.  The run's
        // heads are already in H (block hb0 + l in lane l, block sf in lane
        // hoff).
        const uint32_t hoff = uni((uint32_t)(sf - hb0));
        for (uint32_t lo = vmin;;) {
            const uint32_t hi = (vmax - lo < kWin) ? vmax : lo + (kWin - 1);
            bool any = lo == vmin;  // the first window holds the first candidate
            if (!any) {
#pragma unroll
                for (int i = 0; i < V; ++i) {
                    const uint32_t vi = cand_at(i);
                    any = any || (ln + 64 * i < nlive && vi >= lo && vi <= hi);
                }
            }
            if (__ballot(any)) {
                // blocks b0 .. b1 of the run (relative to sf) overlap [lo, hi]
                const uint32_t c_lo = popc(__ballot(H <= lo)), c_hi = popc(__ballot(H <= hi));
                const uint32_t b0 = c_lo > hoff ? c_lo - 1 - hoff : 0;
                const uint32_t b1 = min(c_hi > hoff ? c_hi - 1 - hoff : 0, r - 1);
                const uint32_t span = hi - lo;
                // the whole bitmap cleared in two 16-byte stores a lane (a narrow
                // window uses fewer words, but a fixed count needs no loop)
                static_assert(kBmWords == 2 * 64 * 4, "two uint4 stores a lane clear the bitmap");
                reinterpret_cast<uint4*>(bm)[ln] = make_uint4(0, 0, 0, 0);
                reinterpret_cast<uint4*>(bm)[64 + ln] = make_uint4(0, 0, 0, 0);
                wave_sync();
                for (uint32_t p0 = b0; p0 <= b1; p0 += NBB) {
                    const uint32_t np = min((uint32_t)NBB, b1 + 1 - p0);
                    SegRec rec[NBB];
                    uint32_t sh[NBB];
                    uint32_t w0[NBB][SEG / 64], w1[NBB][SEG / 64];
                    // every record of the piece first (scalar loads, one wait),
                    // past np a copy of the piece's last block, loaded but not used
#pragma unroll
                    for (int g = 0; g < NBB; ++g) rec[g] = A.ix.rec[R.segnum + sf + min(p0 + (uint32_t)g, b1)];
#pragma unroll
                    for (int g = 0; g < NBB; ++g) {
                        const uint32_t lb = desc_lb(rec[g].desc);
                        const char* q = (const char*)(A.ix.words + R.woff + rec[g].off + 1);
                        sh[g] = ln * lb;
                        // 32-bit byte offsets from the block's (scalar) address, so
                        // that each load takes the scalar base + vector offset form
                        const uint32_t bo = (sh[g] >> 5) * 4u;
#pragma unroll
                        for (int t = 0; t < SEG / 64; ++t) {
                            const uint32_t* w = (const uint32_t*)(q + (bo + 8u * (uint32_t)t * lb));
                            w0[g][t] = w[0];
                            w1[g][t] = w[1];
                        }
                    }
#pragma unroll
                    for (int g = 0; g < NBB; ++g) {
                        if ((uint32_t)g < np) {
                            const uint32_t s = sf + p0 + (uint32_t)g;
                            const uint32_t lb = desc_lb(rec[g].desc), mask = (1u << lb) - 1u;
                            // the block fit's intercept in a vector register once per block, so
                            // that each value's fma reads one scalar operand (the slope) only
                            const float icv = vgpr_copy(rec[g].icpt);
                            const uint32_t add = (uint32_t)rec[g].base - lo;
                            const auto insert = [&](auto tail_c) {
                                constexpr bool TAIL = decltype(tail_c)::value;
                                const uint32_t len = TAIL ? seg_len(R, s, SEG) : (uint32_t)SEG;
#pragma unroll
                                for (int t = 0; t < SEG / 64; ++t) {
                                    const uint32_t j = ln + 64 * t;
                                    const uint32_t f = __builtin_amdgcn_alignbit(w1[g][t], w0[g][t], sh[g]) & mask;
                                    uint32_t x;
                                    if constexpr (CODEC == EXC_SEGLRC)
                                        x = f + (uint32_t)f2i_sat(fmaf((float)j, rec[g].slope, icv)) + add;
                                    else
                                        x = block_value<CODEC, SEG>(R, s, rec[g], j, f) - lo;
                                    if ((!TAIL || j < len) && x <= span)
                                        __hip_atomic_fetch_or(bm + (x & (kBmWords - 1u)), 1u << (x / kBmWords),
                                                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                                }
                            };
                            if (s + 1 == R.nseg)
                                insert(std::true_type{});
                            else
                                insert(std::false_type{});
                        }
                    }
                }
                wave_sync();
#pragma unroll
                for (int i = 0; i < V; ++i) {
                    const uint32_t vi = cand_at(i), x = vi - lo;
                    if (ln + 64 * i < nlive && x <= span)
                        found[i] = (bm[x & (kBmWords - 1u)] >> (x / kBmWords)) & 1u;
                }
                wave_sync();
            }
            if (hi == vmax) break;
            lo = hi + 1;
        }
    } else if (CODEC == EXC_PARAPFD && r <= cmax && ((vmax - vmin) / kWinPfd) < kMaxWindows) {
        // ---- ParaPFD: the run sf .. sf+r-1 decoded in registers straight into
        // the bitmap of each docID window, as the LRC/SegLRC path does (the
        // reference walks one block serially per candidate,
        // GC/ParaPFD/ParaPFD_Intersection_FRAC/ParaPFD_Intersection_kernel.cu:322-455).
        // Each lane takes PV consecutive fields of a block; lanes e < en load
        // exception e with the fields.  A block with exceptions is staged in
        // the wave's distinct-segment area (idle on this path), patched there
        // lane-parallel and read back; then the block's head replaces its
        // first gap and a wave scan of the lanes' sums turns the gaps into
        // docIDs.  No candidate searches: each tests its bit.
        constexpr int PV = SEG / 64;
        constexpr int NBP = 4;  // blocks per load round
        uint32_t* bm = W.buf;                     // kBufWords - SEG words of bitmap
        uint32_t* stg = W.buf + (kBufWords - SEG);  // then one block's staging
        const SegRec gr = A.ix.rec[R.segnum + sf + min(ln, r - 1)];
        const uint32_t hd = ln < r ? gr.head : kNone;
        for (uint32_t lo = vmin;;) {
            const uint32_t hi = (vmax - lo < kWinPfd) ? vmax : lo + (kWinPfd - 1);
            bool any = false;
#pragma unroll
            for (int i = 0; i < V; ++i) {
                const uint32_t vi = cand_at(i);
                any = any || (ln + 64 * i < nlive && vi >= lo && vi <= hi);
            }
            if (__ballot(any)) {
                const uint32_t c_lo = popc(__ballot(hd <= lo)), c_hi = popc(__ballot(hd <= hi));
                const uint32_t b0 = c_lo ? c_lo - 1 : 0, b1 = c_hi ? c_hi - 1 : 0;
                const uint32_t nw = ((hi - lo) >> 5) + 1;
                for (uint32_t e = ln; e < nw; e += 64) bm[e] = 0;
                wave_sync();
                for (uint32_t p0 = b0; p0 <= b1; p0 += NBP) {
                    const uint32_t np = min((uint32_t)NBP, b1 + 1 - p0);
                    uint32_t desc[NBP], head[NBP], len[NBP];
                    ExcPrefetch xp[NBP];
                    const uint32_t* data[NBP];
                    uint32_t f[NBP][PV];
#pragma unroll
                    for (int g = 0; g < NBP; ++g) {
                        // past np: a copy of the piece's last block, loaded but not used
                        const int src = (int)min(p0 + g, b1);
                        desc[g] = uni(__builtin_amdgcn_readlane(gr.desc, src));
                        head[g] = uni(__builtin_amdgcn_readlane(gr.head, src));
                        len[g] = seg_len(R, sf + (uint32_t)src, SEG);
                        data[g] = A.ix.words + R.woff + uni(__builtin_amdgcn_readlane(gr.off, src)) + 1;
                        load_fields<CODEC, SEG, false>(data[g], desc_lb(desc[g]), len[g], f[g]);
                        xp[g] = prefetch_exceptions<CODEC>(data[g], desc[g], len[g]);
                    }
#pragma unroll
                    for (int g = 0; g < NBP; ++g) {
                        if ((uint32_t)g < np) {
                            if (desc_en(desc[g])) {
                                // stage, patch, read back (each lane stages and
                                // reads its own words; the barrier after the
                                // staging also orders the previous block's reads
                                // before this block's patch)
#pragma unroll
                                for (int t = 0; t < PV; ++t) stg[ln * PV + t] = f[g][t];
                                wave_sync();
                                pfd_patch_lds<true>(data[g], desc[g], len[g], xp[g], stg);
                                wave_sync();
#pragma unroll
                                for (int t = 0; t < PV; ++t) f[g][t] = stg[ln * PV + t];
                            }
                            if (ln == 0) f[g][0] = head[g];
                            uint32_t acc = 0;
#pragma unroll
                            for (int t = 0; t < PV; ++t) { acc += f[g][t]; f[g][t] = acc; }
                            const uint32_t base = wave_incl_scan(acc) - acc - lo;
#pragma unroll
                            for (int t = 0; t < PV; ++t) {
                                const uint32_t j = ln * PV + t, x = f[g][t] + base;
                                if (j < len[g] && x <= hi - lo)
                                    __hip_atomic_fetch_or(bm + word_of_bit(x), 1u << (x & 31), __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_WAVEFRONT);
                            }
                        }
                    }
                }
                wave_sync();
#pragma unroll
                for (int i = 0; i < V; ++i) {
                    const uint32_t vi = cand_at(i), x = vi - lo;
                    if (ln + 64 * i < nlive && vi >= lo && x <= hi - lo) found[i] = (bm[x >> 5] >> (x & 31)) & 1u;
                }
                wave_sync();
            }
            if (hi == vmax) break;
            lo = hi + 1;
        }
    } else if (CODEC == EXC_PARAPFD && r <= cmax) {
        // ---- the run of segments sf .. sf+r-1 decoded piece by piece into LDS;
        // one gather brings every segment record of the run, one load round
        // per piece brings the packed words, sorted candidates resolve in the
        // piece that reaches their value
        constexpr int P = (CODEC == EXC_PARAPFD) ? (int)(512 / SEG) : NB;
        uint32_t v[V];
#pragma unroll
        for (int i = 0; i < V; ++i) v[i] = cand_at(i);
        const uint32_t gsel = min(ln, r - 1);
        const SegRec gr = A.ix.rec[R.segnum + sf + gsel];
        bool done[V];
#pragma unroll
        for (int i = 0; i < V; ++i) done[i] = ln + 64 * i >= nlive;
        for (uint32_t p0 = 0; p0 < r; p0 += P) {
            const uint32_t np = min((uint32_t)P, r - p0);
            SegRec rec[P];
            uint32_t len[P];
            const uint32_t* data[P];
            uint32_t f[P][SEG / 64];
#pragma unroll
            for (int g = 0; g < P; ++g) {
                {  // past np: a copy of the piece's last block, loaded but not used
                    const int src = (int)min(p0 + g, r - 1);
                    rec[g].off = uni(__builtin_amdgcn_readlane(gr.off, src));
                    rec[g].desc = uni(__builtin_amdgcn_readlane(gr.desc, src));
                    rec[g].slope = __uint_as_float(uni(__builtin_amdgcn_readlane(__float_as_uint(gr.slope), src)));
                    rec[g].icpt = __uint_as_float(uni(__builtin_amdgcn_readlane(__float_as_uint(gr.icpt), src)));
                    rec[g].base = (int32_t)uni(__builtin_amdgcn_readlane((uint32_t)gr.base, src));
                    rec[g].head = uni(__builtin_amdgcn_readlane(gr.head, src));
                    len[g] = seg_len(R, sf + (uint32_t)src, SEG);
                    data[g] = A.ix.words + R.woff + rec[g].off + 1;
                    load_fields<CODEC, SEG>(data[g], desc_lb(rec[g].desc), len[g], f[g]);
                }
            }
#pragma unroll
            for (int g = 0; g < P; ++g)
                if ((uint32_t)g < np)
                    finish_block<CODEC, SEG>(R, sf + p0 + g, rec[g], len[g], data[g], f[g], W.buf + g * SEG, nullptr);
            const uint32_t total = (np - 1) * SEG + seg_len(R, sf + p0 + np - 1, SEG);
            wave_sync();
            const uint32_t last = uni(W.buf[total - 1]);
            bool any = false;
#pragma unroll
            for (int i = 0; i < V; ++i) {
                if (!done[i] && v[i] <= last) {
                    found[i] = lds_contains(W.buf, total, v[i]);
                    done[i] = true;
                }
                any = any || !done[i];
            }
            wave_sync();
            if (!__ballot(any)) break;
        }
    } else {
        // beyond 64 heads: every stride-th head of the range in one load round
        // (lane l: head l * stride), shared by all candidate rows; a candidate
        // then binary-searches only the stride heads between two samples
        const uint32_t stride = small ? 1u : uni((uint32_t)((rng + 63) >> 6));
        const uint32_t G = small ? H : (ln * stride < (uint32_t)rng ? head[hb0 + ln * stride] : kNone);
        const uint32_t h63 = uni(__builtin_amdgcn_readlane(G, 63));
        // segment of a candidate (-1: precedes the list, -2: no candidate)
        const auto seg_of = [&](uint32_t vi, bool live) -> int32_t {
            int64_t cnt;
            if (small) {
                cnt = lane_count_le(H, h63, vi);
            } else {
                const uint32_t c = lane_count_le(G, h63, vi);  // samples <= vi
                uint32_t lo = c ? (c - 1) * stride + 1 : 0u, hi = c ? min(c * stride, (uint32_t)rng) : 0u;
                if (live)
                    while (lo < hi) {
                        const uint32_t m = (lo + hi) >> 1;
                        if (head[hb0 + m] <= vi) lo = m + 1; else hi = m;
                    }
                cnt = lo;
            }
            return live ? (int32_t)(hb0 + cnt - 1) : -2;
        };
        if constexpr (CODEC == EXC_SEGLRC) {
            // ---- one lane per candidate, row by row (a loop, so its state stays
            // small): the segment record, then the words of the positions the
            // block fit allows ([p - rl - 1, p + rr + 1], up to kWinWords
            // words) in one load round, staged in LDS (word k of lane l at
            // buf[64k + l]) and binary-searched there.  A candidate outside the
            // window's values, or a window too wide, falls back to the exact
            // search in HBM (plain binary search when the search mode is binary).
            uint32_t fmask = 0;
            for (int i = 0; i < V; ++i) {
                const uint32_t e = ln + 64u * (uint32_t)i;
                if (!__ballot(e < nlive)) break;  // rows fill in order
                const uint32_t vi = cand_at(i);
                const int32_t sc = seg_of(vi, e < nlive);
                if (!__ballot(sc >= 0)) continue;
                const uint32_t si = (uint32_t)(sc >= 0 ? sc : (int32_t)sf);
                const SegRec rc = A.ix.rec[R.segnum + si];
                const uint32_t len = seg_len(R, si, SEG);
                const uint32_t* data = A.ix.words + R.woff + rc.off + 1;
                const bool f = win_search(W.buf, data, len, rc, vi, sc >= 0, search);
                if (f) fmask |= 1u << i;
            }
#pragma unroll
            for (int i = 0; i < V; ++i) found[i] = (fmask >> i) & 1u;
        } else {
        uint32_t v[V];
#pragma unroll
        for (int i = 0; i < V; ++i) v[i] = cand_at(i);
        int32_t s[V];
#pragma unroll
        for (int i = 0; i < V; ++i) s[i] = seg_of(v[i], ln + 64 * i < nlive);
        if constexpr (CODEC == EXC_PARAPFD) {
            // distinct segments of the candidates (candidate-order runs)
            uint64_t fm[V];
            uint32_t nd = 0;
#pragma unroll
            for (int i = 0; i < V; ++i) {
                // segment of the previous candidate: lane - 1 (a DPP wave_shr:1
                // move), or lane 63 of the previous row
                int32_t sp = __builtin_amdgcn_update_dpp(0, s[i], 0x138, 0xf, 0xf, false);
                if (ln == 0) sp = (i > 0) ? (int32_t)__builtin_amdgcn_readlane((uint32_t)s[i > 0 ? i - 1 : 0], 63) : -2;
                fm[i] = __ballot(s[i] >= 0 && s[i] != sp);
                nd += popc(fm[i]);
            }
            nd = uni(nd);
            bool sparse = false;
            if constexpr (SEG == 128) sparse = (A.ix.flags & kIxOctants) && nlive <= kPfdSparse * nd;
            if (sparse) {
                if constexpr (SEG == 128) {
                // ---- sparse: few candidates per block.  Each candidate's
                // eighth of a block, the one its value falls in by the block's
                // values at the eighth boundaries (the head; the median of
                // .segMedian, the reference's half-select,
                // GC/ParaPFD/ParaPFD_Intersection_FRAC/ParaPFD_Intersection_kernel.cu:322-455;
                // the values at the quarter positions in the record and at
                // the quarters' midpoints in pfd_oct), is decoded by one row
                // of 4 lanes, 4 positions a lane: fields, exceptions ORed
                // into the row's LDS stage, prefix from the eighth's first
                // value, compare.  Sixteen candidates a pass; their segments
                // wait in the buffer's first kChunk words.
                constexpr uint32_t EP = 16, PL = EP / 4;
                uint32_t* segs = W.buf;
                uint32_t* stage = W.buf + kChunk;
                static_assert(kChunk + 16 * EP <= kBufWords, "segments and sixteen stages fit the buffer");
#pragma unroll
                for (int i = 0; i < V; ++i) segs[ln + 64 * i] = (uint32_t)s[i];
                wave_sync();
                const uint32_t r = ln >> 2, k = ln & 3u;
                uint32_t hitm = 0;  // bit i: this lane's row-i candidate was found
                for (uint32_t c0 = 0; c0 < nlive; c0 += 16) {
                    const uint32_t c = c0 + r;
                    const int32_t sc = c < nlive ? (int32_t)segs[c] : -1;
                    const bool ok = sc >= 0;
                    const uint32_t vc = ok ? cand[c] : 0u;
                    const uint32_t si = ok ? (uint32_t)sc : sf;
                    const SegRec rc = A.ix.rec[R.segnum + si];
                    const uint4 oc = reinterpret_cast<const uint4*>(A.ix.pfd_oct)[R.segnum + si];
                    const uint32_t len = seg_len(R, si, SEG);
                    // quarter, then its half
                    const bool up = vc >= rc.rl;
                    const bool q_hi = up ? vc >= rc.rr : vc >= (uint32_t)rc.base;
                    const uint32_t q = (up ? 2u : 0u) + (q_hi ? 1u : 0u);
                    const uint32_t qfirst = up ? (q_hi ? rc.rr : rc.rl) : (q_hi ? (uint32_t)rc.base : rc.head);
                    const uint32_t mid = q == 0 ? oc.x : q == 1 ? oc.y : q == 2 ? oc.z : oc.w;
                    const bool h_hi = vc >= mid;
                    const uint32_t qs = pfd_qb(len, q), qe = pfd_qb(len, q + 1), qm = qs + ((qe - qs) >> 1);
                    const uint32_t p0 = h_hi ? qm : qs, pe = h_hi ? qe : qm;
                    const uint32_t pn = ok ? pe - p0 : 0u;
                    const uint32_t first = h_hi ? mid : qfirst;
                    const uint32_t* data = A.ix.words + R.woff + rc.off + 1;
                    const uint32_t lb = desc_lb(rc.desc), en = ok ? desc_en(rc.desc) : 0u;
                    uint32_t f[PL];
#pragma unroll
                    for (uint32_t t = 0; t < PL; ++t) f[t] = bits_at(data, lb, p0 + k * PL + t);
                    if (__ballot(en != 0)) {
                        uint32_t* st = stage + r * EP;
#pragma unroll
                        for (uint32_t t = 0; t < PL; ++t) st[k * PL + t] = f[t];
                        wave_sync();
                        const uint32_t eb = desc_eb(rc.desc), hb = desc_hb(rc.desc);
                        const uint32_t* pp = data + words_for(len, lb);
                        const uint32_t* hp = pp + words_for(en, eb);
                        for (uint32_t e = k; e < en; e += 4) {
                            const uint32_t qq = bits_at(pp, eb, e) - p0;
                            if (qq < pn)
                                __hip_atomic_fetch_or(st + qq, bits_at(hp, hb, e) << lb, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_WAVEFRONT);
                        }
                        wave_sync();
#pragma unroll
                        for (uint32_t t = 0; t < PL; ++t) f[t] = st[k * PL + t];
                        wave_sync();
                    }
                    // values: the eighth's first value, then its gaps (its
                    // first gap stands for that value); the lanes' sums
                    // scanned within each row of 4 by DPP row shifts
                    if (k == 0) f[0] = 0;
#pragma unroll
                    for (uint32_t t = 1; t < PL; ++t) f[t] += f[t - 1];
                    uint32_t x = f[PL - 1], y;
                    y = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, true);
                    x += k >= 1 ? y : 0u;
                    y = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, true);
                    x += k >= 2 ? y : 0u;
                    const uint32_t off = x - f[PL - 1] + first;
                    bool hit = false;
#pragma unroll
                    for (uint32_t t = 0; t < PL; ++t) hit = hit || (k * PL + t < pn && f[t] + off == vc);
                    const uint64_t m = __ballot(hit);
                    uint32_t hit16 = 0;
#pragma unroll
                    for (uint32_t g = 0; g < 16; ++g) hit16 |= ((m >> (4 * g)) & 0xfull) ? 1u << g : 0u;
                    // candidates c0 .. c0 + 15 belong to lanes c0 % 64 .. + 15 of row c0 / 64
                    if ((ln >> 4) == ((c0 & 63u) >> 4)) hitm |= ((hit16 >> (ln & 15u)) & 1u) << (c0 >> 6);
                }
#pragma unroll
                for (int i = 0; i < V; ++i) found[i] = (hitm >> i) & 1u;
                wave_sync();
                }
            } else {
            // ---- decode each distinct needed segment once, ND per round into
            // the buffer's head, the list of segments after the decode area:
            // four a round when the list fits the buffer's last 256 words,
            // else two (the list taking kChunk words)
            uint32_t di[V];
            const auto rounds = [&](auto nd_c, uint32_t* dl) {
                constexpr int ND = decltype(nd_c)::value;
                static_assert(ND >= 1, "one segment fits the decode area");
                uint32_t before = 0;
#pragma unroll
                for (int i = 0; i < V; ++i) {
                    const bool first = (fm[i] >> ln) & 1ull;
                    di[i] = before + below_me(fm[i]) + (first ? 1u : 0u) - 1u;
                    if (first) dl[di[i]] = (uint32_t)s[i];
                    before += popc(fm[i]);
                }
                wave_sync();
                for (uint32_t r0 = 0; r0 < nd; r0 += ND) {
                    const uint32_t gn = min((uint32_t)ND, nd - r0);
                    decode_blocks<CODEC, SEG, ND, true>(A.ix, R, gn, [&](int g) { return dl[r0 + g]; }, W.buf);
                    wave_sync();
#pragma unroll
                    for (int i = 0; i < V; ++i) {
                        if (s[i] >= 0 && di[i] >= r0 && di[i] < r0 + ND)
                            found[i] = lds_contains(W.buf + (di[i] - r0) * SEG, seg_len(R, (uint32_t)s[i], SEG), v[i]);
                    }
                    wave_sync();
                }
            };
            constexpr uint32_t kWideDec = 4 * SEG;  // four segments, then a list of at most kBufWords - 4 SEG
            if constexpr (kWideDec + 64 <= kBufWords) {
                if (nd <= kBufWords - kWideDec)
                    rounds(std::integral_constant<int, 4>{}, W.buf + kWideDec);
                else
                    rounds(std::integral_constant<int, (int)(kPfdDecWords / SEG)>{}, W.dlist);
            } else {
                rounds(std::integral_constant<int, (int)(kPfdDecWords / SEG)>{}, W.dlist);
            }
            }
        } else {
            // LRC: binary search inside the segment
#pragma unroll
            for (int i = 0; i < V; ++i) {
                if (s[i] >= 0) {
                    const SegRec rc = A.ix.rec[R.segnum + s[i]];
                    found[i] = lrc_bsearch(A.ix.words + R.woff + rc.off + 1, desc_lb(rc.desc), (uint32_t)s[i] * SEG,
                                           seg_len(R, (uint32_t)s[i], SEG), R, v[i]);
                }
            }
        }
        }
    }

    // ---- order-preserving compaction of survivors, row by row: a row is read
    // before it is written and its survivors move only to lower positions,
    // so later rows are intact when read
    uint32_t cnt = 0;
    wave_sync();
#pragma unroll
    for (int i = 0; i < V; ++i) {
        const uint64_t m = __ballot(found[i]);
        const uint32_t vi = cand_at(i);
        // lanes without a survivor write into the buffer behind the candidates
        // (free between lists) instead of branching round the store
        cand[found[i] ? cnt + below_me(m) : kChunk + ln] = vi;
        cnt += popc(m);
    }
    wave_sync();
    return uni(cnt);
}

// ------------------------------------------------------------ uncompressed

// First index in [lo, hi) of the sorted array L with L[idx] >= v (hi if none):
// 64-ary narrowing by the whole wave, one load round per factor of 64.
__device__ __forceinline__ uint32_t wave_lower_bound(const uint32_t* __restrict__ L, uint32_t lo, uint32_t hi, uint32_t v)
{
    const uint32_t ln = lane_id();
    while (hi - lo > 64) {
        const uint32_t step = (hi - lo + 63) >> 6;  // 64 samples cover [lo, hi)
        const uint32_t p = lo + ln * step;
        const uint32_t c = popc(__ballot(p < hi && L[p] < v));  // samples below v: a prefix
        if (c == 0) return lo;
        const uint32_t nlo = lo + (c - 1) * step + 1, nhi = min(hi, lo + c * step);
        lo = uni(nlo);
        hi = uni(nhi);
    }
    const uint32_t p = lo + ln;
    return lo + popc(__ballot(p < hi && L[p] < v));
}

// Is v in the sorted L[a .. b)?  interp: interpolation search
// (GU/cuda_interpolationSearch/cuda_interpolation_kernel.cu:254-297), with
// the probe position clamped to [left, right] (the reference's float
// interpolation can land one past right on long lists, where its loop stops
// progressing) and each step loading the probe and both of its neighbours in
// one round; otherwise the plain binary search of mqSearch
// (GU/cuda_binarySearch/cuda_binarySearch_kernel.cu:258-293).  Reads stay in
// L[a .. b).
__device__ __forceinline__ bool raw_find(const uint32_t* __restrict__ L, uint32_t a, uint32_t b, uint32_t t,
                                         uint32_t r, uint32_t v, bool interp)
{
    if (a >= b || v < t || v > r) return false;
    int32_t lo = (int32_t)a, hi = (int32_t)b - 1;
    if (!interp) {
        while (lo <= hi) {
            const int32_t m = (lo + hi) >> 1;
            const uint32_t q = L[m];
            if (q == v) return true;
            if (q < v) lo = m + 1; else hi = m - 1;
        }
        return false;
    }
    while (lo <= hi && t <= v && v <= r) {
        int32_t m = lo;
        if (r != t) {
            const float fr = (float)(hi - lo) * (float)(v - t) / (float)(r - t);
            m = min(hi, max(lo, lo + (int32_t)fr));
        }
        const uint32_t q = L[m];
        const uint32_t qn = m < hi ? L[m + 1] : 0u, qp = m > lo ? L[m - 1] : 0u;
        if (q == v) return true;
        if (q < v) {
            lo = m + 1;
            t = qn;
        } else {
            hi = m - 1;
            r = qp;
        }
    }
    return false;
}

// Positions [lo, hi) where docID v can sit in uncompressed list R: from the
// list's regression line, x = (v - fIntercept) / fSlope with every element i
// inside x - nRangeLeft < i < x + nRangeRight
// (GU/cuda_runRegressionAlgorithm/cuda_linearRegression_kernel.cu:262-311,
// which takes the window once per block of the shortest list), or from its
// learned model, prediction +- recorded max error + 2.  Either is a guess
// the caller verifies: float rounding may move an edge.
__device__ __forceinline__ void raw_window(const DevIndex& ix, const TermRec& R, uint32_t v, bool learned,
                                           int64_t& lo, int64_t& hi)
{
    float x, l, r;
    if (learned) {
        const float* prm = ix.mparams + (uint64_t)R.model * model_stride(R.mkind);
        x = model_pred(R.mkind, prm, R.mx0, R.mxs, R.n, v);
        l = r = R.merr + 2.0f;
    } else {
        x = ((float)v - __int_as_float(R.icpt)) / R.slope;
        l = R.range_l;
        r = R.range_r;
    }
    // NaN and out-of-range guesses clamp to [-1, 4.3e9]: an empty window at worst
    lo = (int64_t)floorf(fminf(fmaxf(x - l, -1.0f), 4.3e9f));
    hi = (int64_t)ceilf(fminf(fmaxf(x + r, -1.0f), 4.3e9f)) + 1;
}

// x clamped into [lo, hi]
__device__ __forceinline__ uint32_t clamp_pos(int64_t x, uint32_t lo, uint32_t hi)
{
    return x < (int64_t)lo ? lo : x > (int64_t)hi ? hi : (uint32_t)x;
}

// Uncompressed list R filters the unit's candidates: the wave bounds the
// part [a, b) of R that can hold [vmin, vmax] (inside the regression or
// learned window of vmin .. vmax when the list's elements around it show
// that the window holds them all -- the reference's per-block borders), then
// each lane searches [a, b) for its candidates, row by row, by binary or
// interpolation search; order-preserving compaction as for the codecs.
// (A window per candidate does not pay here: its lookup costs what the top
// levels of the binary search it would skip cost, and those are cache hits
// shared by the wave.)
__device__ uint32_t probe_raw(const ProbeArgs& A, const TermRec& R, const LdsWave& W, uint32_t nlive, uint32_t knobs)
{
    constexpr int V = (int)(kChunk / 64);
    const uint32_t ln = lane_id();
    const uint32_t* L = A.ix.words + R.woff;
    const uint32_t n = R.n;
    const uint32_t vmin = uni(W.cand[0]), vmax = uni(W.cand[nlive - 1]);
    const uint32_t search = knob_search(knobs);
    const bool learned = search == SEARCH_LEARNED && R.model != kNoModel;
    const bool windowed = learned || (search != SEARCH_BINARY && search != SEARCH_INTERP && R.range_l < 1.0e9f);
    uint32_t lo = 0, hi = n;
    if (windowed) {
        int64_t l0, h0, l1, h1;
        raw_window(A.ix, R, vmin, learned, l0, h0);
        raw_window(A.ix, R, vmax, learned, l1, h1);
        const uint32_t wl = uni(clamp_pos(l0, 0, n));
        const uint32_t wh = uni(clamp_pos(h1, wl, n));
        if ((wl == 0 || L[wl - 1] < vmin) && (wh == n || L[wh] > vmax)) {
            lo = wl;
            hi = wh;
        }
    }
    const uint32_t a = wave_lower_bound(L, lo, hi, vmin);
    const uint32_t b = a < hi ? wave_lower_bound(L, a, hi, vmax + 1u) : a;
    if (a >= b) return 0;
    const uint32_t t = uni(L[a]), r = uni(L[b - 1]);
    const bool interp = search == SEARCH_INTERP;
    bool found[V];
#pragma unroll
    for (int i = 0; i < V; ++i) {
        const uint32_t e = ln + 64 * i;
        found[i] = e < nlive && raw_find(L, a, b, t, r, W.cand[e], interp);
    }
    uint32_t cnt = 0;
    wave_sync();
#pragma unroll
    for (int i = 0; i < V; ++i) {
        const uint64_t m = __ballot(found[i]);
        const uint32_t e = ln + 64 * i;
        const uint32_t vi = e < nlive ? W.cand[e] : kNone;
        if (found[i]) W.cand[cnt + below_me(m)] = vi;
        cnt += popc(m);
    }
    wave_sync();
    return uni(cnt);
}

// Uncompressed list R in hash buckets on its top docID bits
// (GU/cuda_hashSegment_new/cuda_hashSegment_kernel.cu:269-307): a candidate v
// searches only bucket b = v >> shift, [start_b, start_(b+1)) of R.  The
// reference takes b unchecked, so a docID past the list's last bucket reads
// beyond its table; here b is clamped to the last bucket, which is exact
// (the validated table brackets its docIDs: every element of a bucket below
// b is below b << shift, so a docID past the last bucket can only be in it).
// Binary search (EXC_OPT_SEARCH 0, the reference's) or interpolation search
// inside the bucket; reads stay inside the bucket.
__device__ uint32_t probe_raw_hs(const ProbeArgs& A, const TermRec& R, const LdsWave& W, uint32_t nlive, uint32_t knobs)
{
    constexpr int V = (int)(kChunk / 64);
    const uint32_t ln = lane_id();
    const uint32_t* L = A.ix.words + R.woff;
    const uint32_t* starts = A.ix.hash + R.aux;
    const uint32_t shift = (uint32_t)R.icpt, lastb = R.nseg - 1;
    const bool interp = knob_search(knobs) != SEARCH_BINARY;
    uint32_t fmask = 0;
    for (int i = 0; i < V; ++i) {
        const uint32_t e = ln + 64u * (uint32_t)i;
        if (!__ballot(e < nlive)) break;  // rows fill in order
        if (e < nlive) {
            const uint32_t v = W.cand[e];
            const uint32_t b = min(v >> shift, lastb);
            const uint32_t lo = starts[b], hi = starts[b + 1];
            if (lo < hi) {
                const uint32_t t = interp ? L[lo] : 0u, r = interp ? L[hi - 1] : 0xffffffffu;
                if (raw_find(L, lo, hi, t, r, v, interp)) fmask |= 1u << i;
            }
        }
    }
    uint32_t cnt = 0;
    wave_sync();
#pragma unroll
    for (int i = 0; i < V; ++i) {
        const bool fi = (fmask >> i) & 1u;
        const uint64_t m = __ballot(fi);
        const uint32_t e = ln + 64 * i;
        const uint32_t vi = e < nlive ? W.cand[e] : kNone;
        if (fi) W.cand[cnt + below_me(m)] = vi;
        cnt += popc(m);
    }
    wave_sync();
    return uni(cnt);
}

// HS_SegLRC: decode elements [k*CH, k*CH+CH) of list R into dst; returns count.
// Record of each element = last record whose first element is <= its index
// (HS_SegLRC_Intersection_kernel.cu:274-327, over the non-empty buckets only).
template <int CH>
__device__ uint32_t decode_hs_chunk(const DevIndex& ix, const TermRec& R, uint32_t k, uint32_t* dst)
{
    constexpr int V = CH / 64;
    const uint32_t ln = lane_id();
    const uint32_t e0 = k * CH;
    const uint32_t clen = uni(min((uint32_t)CH, R.n - e0));
    const uint32_t* heads = ix.seg_head + R.segnum;
    const uint32_t nr = R.nseg;
    const uint32_t r0 = uni(wave_upper_bound(heads, nr, e0) - 1);
    const uint32_t hreg = (r0 + 1 + ln < nr) ? heads[r0 + 1 + ln] : kNone;
    const uint32_t h63 = uni(__builtin_amdgcn_readlane(hreg, 63));
#pragma unroll
    for (int i = 0; i < V; ++i) {
        const uint32_t j = ln + 64 * i;
        const uint32_t e = e0 + j;
        const uint32_t pos = lane_count_le(hreg, h63, e);
        if (j < clen) {
            uint32_t b = r0 + pos;
            if (pos == 64) {
                uint32_t lo = r0 + 65, hi = nr;
                while (lo < hi) {
                    const uint32_t m = (lo + hi) >> 1;
                    if (heads[m] <= e) lo = m + 1; else hi = m;
                }
                b = lo - 1;
            }
            const SegRec rc = ix.rec[R.segnum + b];
            const uint32_t* data = ix.words + R.woff + rc.off + 1;
            dst[j] = seg_val(data, desc_lb(rc.desc), e - rc.head, rc);
        }
    }
    return clen;
}

// Record of the bucket holding docID v in an HS_SegLRC list, and the number
// of elements in it; false when v is past the last bucket or its bucket is
// empty.  A dense list's record is its bucket; otherwise the occupancy pair of
// the bucket's group of 32 gives presence and rank.
struct HsList {
    const SegRec* rec;
    const uint32_t* heads;
    const uint2* occ;  // null for a dense list
    uint32_t shift, nb, nr, n;
};

__device__ __forceinline__ HsList hs_list(const DevIndex& ix, const TermRec& R)
{
    HsList h;
    h.rec = ix.rec + R.segnum;
    h.heads = ix.seg_head + R.segnum;
    h.shift = (uint32_t)R.icpt;
    h.nr = R.nseg;
    h.n = R.n;
    h.occ = nullptr;
    h.nb = h.nr;
    const uint32_t aux = uni(R.aux);
    if (aux != kDenseBuckets) {
        h.occ = reinterpret_cast<const uint2*>(ix.hash + aux + 2);
        h.nb = uni(ix.hash[aux]);
    }
    return h;
}

__device__ __forceinline__ bool hs_record(const HsList& h, uint32_t v, uint32_t& r)
{
    const uint32_t b = v >> h.shift;
    if (b >= h.nb) return false;
    r = b;
    if (h.occ) {
        const uint2 p = h.occ[b >> 5];
        const uint32_t bit = b & 31u;
        if (!((p.x >> bit) & 1u)) return false;
        r = p.y + popc(p.x & ((1u << bit) - 1u));
    }
    return true;
}

__device__ __forceinline__ uint32_t hs_len(const HsList& h, uint32_t r, uint32_t head)
{
    return (r + 1 < h.nr ? h.heads[r + 1] : h.n) - head;
}

// HS_SegLRC probe: O(1) bucket, fit window, exact fallback
// (HS_SegLRC_Intersection_kernel.cu:334-401 plus the fallback it lacks).
__device__ uint32_t probe_hs(const ProbeArgs& A, const TermRec& R, const LdsWave& W, uint32_t nlive, uint32_t knobs)
{
    const uint32_t search = knob_search(knobs);
    constexpr int V = kV;
    const uint32_t ln = lane_id();
    const HsList H = hs_list(A.ix, R);
    const uint32_t nb = H.nb;
    uint32_t fmask = 0;
    if ((uint64_t)R.n > 512ull * nb) {
        // large buckets (fixed shifts on long lists): their fits are too loose
        // for a one-load window; two candidates per lane in flight, each
        // predicting its position and walking / binary-searching in HBM
#pragma unroll
        for (int i = 0; i < V; i += 2) {
            SegRec rec[2];
            uint32_t len[2], vv[2];
            bool ok[2];
#pragma unroll
            for (int x = 0; x < 2; ++x) {
                const uint32_t e = ln + 64 * (i + x);
                vv[x] = e < nlive ? W.cand[e] : kNone;
                uint32_t r = 0;
                ok[x] = e < nlive && hs_record(H, vv[x], r);
                len[x] = 0;
                if (ok[x]) {
                    rec[x] = H.rec[r];
                    len[x] = hs_len(H, r, rec[x].head);
                }
            }
#pragma unroll
            for (int x = 0; x < 2; ++x)
                if (ok[x] && len[x] > 0 &&
                    seg_find(A.ix.words + R.woff + rec[x].off + 1, len[x], rec[x], vv[x], search != SEARCH_BINARY))
                    fmask |= 1u << (i + x);
        }
    } else for (int i = 0; i < V; ++i) {
        // buckets of ~256 docIDs: row by row, bucket, record, windowed search
        const uint32_t e = ln + 64u * (uint32_t)i;
        if (!__ballot(e < nlive)) break;  // rows fill in order
        const uint32_t vi = e < nlive ? W.cand[e] : kNone;
        uint32_t r = 0;
        const bool inb = e < nlive && hs_record(H, vi, r);
        if (!__ballot(inb)) continue;
        const SegRec rc = H.rec[r];
        const uint32_t len = inb ? hs_len(H, r, rc.head) : 0u;
        const bool f = win_search(W.buf, A.ix.words + R.woff + rc.off + 1, len, rc, vi, inb && len > 0, search);
        if (f) fmask |= 1u << i;
    }
    // order-preserving compaction, row by row (rows read before written)
    uint32_t cnt = 0;
    wave_sync();
#pragma unroll
    for (int i = 0; i < V; ++i) {
        const bool fi = (fmask >> i) & 1u;
        const uint64_t m = __ballot(fi);
        const uint32_t e = ln + 64 * i;
        const uint32_t vi = e < nlive ? W.cand[e] : kNone;
        if (fi) W.cand[cnt + below_me(m)] = vi;
        cnt += popc(m);
    }
    wave_sync();
    return uni(cnt);
}

// uncompressed lists: .ind2 holds the docIDs themselves
constexpr bool raw_codec(int codec) { return codec == EXC_RAW || codec == EXC_RAW_HS; }

template <int CODEC>
constexpr int wave_lds_words()
{
    return raw_codec(CODEC) ? (int)kChunk : (int)(kChunk + kBufWords);
}

template <int CODEC, int SEG>
__device__ void run_unit(const ProbeArgs& A, uint32_t u, uint32_t* L)
{
    constexpr uint32_t U = kChunk / SEG;  // shortest-list segments per unit
    const LdsWave W{L, L + kChunk, L + kChunk + kPfdDecWords};
    const uint32_t ln = lane_id();
    const UnitRec ur = A.units[u];
    const uint32_t k = uni(ur.k), tc = uni((uint32_t)ur.tc), knobs = uni((uint32_t)ur.knobs);
    const uint32_t l1_first = uni(ur.l1_first), l1_count = uni(ur.l1_count);
    const TermRec* qr = A.qrecs + uni(ur.toff);
    const TermRec R0 = qr[0];
    // the second list's heads over the prologue's blocks, requested before the
    // shortest list is decoded so that their load overlaps the decode
    uint32_t pre_heads = kNone;
    if constexpr (!raw_codec(CODEC) && CODEC != EXC_HS_SEGLRC) {
        if (tc > 1 && l1_count != kUnset && l1_count <= 64 && ln < l1_count)
            pre_heads = A.ix.seg_head[qr[1].segnum + l1_first + ln];
    }
    uint32_t nlive;
    if constexpr (CODEC == EXC_HS_SEGLRC) {
        nlive = decode_hs_chunk<kChunk>(A.ix, R0, k, W.cand);
    } else if constexpr (raw_codec(CODEC)) {
        const uint32_t b0 = k * kChunk;
        nlive = min(kChunk, R0.n - b0);
        const uint32_t* src = A.ix.words + R0.woff + b0;
        for (uint32_t e = ln; e < nlive; e += 64) W.cand[e] = src[e];
    } else {
        const uint32_t s0 = k * U;
        const uint32_t nu = uni(min(U, R0.nseg - s0));
        if constexpr (CODEC == EXC_PARAPFD)
            decode_pfd_phased<SEG, (int)U>(A.ix, R0, s0, nu, W.cand);
        else
            decode_ra_padded<CODEC, SEG, (int)U>(A.ix, R0, s0, nu, W.cand);
        nlive = (nu - 1) * SEG + seg_len(R0, s0 + nu - 1, SEG);
    }
    nlive = uni(nlive);
    wave_sync();
    for (uint32_t j = 1; j < tc && nlive; ++j) {
        const TermRec R = qr[j];
        if constexpr (CODEC == EXC_HS_SEGLRC)
            nlive = probe_hs(A, R, W, nlive, knobs);
        else if constexpr (CODEC == EXC_RAW)
            nlive = probe_raw(A, R, W, nlive, knobs);
        else if constexpr (CODEC == EXC_RAW_HS)
            nlive = probe_raw_hs(A, R, W, nlive, knobs);
        else
            nlive = probe_seg<CODEC, SEG>(A, R, W, nlive, knobs, j == 1 ? l1_first : 0u, j == 1 ? l1_count : kUnset,
                                              pre_heads);
    }
    uint32_t* out = A.stage + uni(ur.soff);
    for (uint32_t e = ln; e < nlive; e += 64) out[e] = W.cand[e];
    if (ln == 0) A.ucount[uni(ur.slot)] = nlive;
    wave_sync();
}

// Persistent work queue.  One returning device-scope atomic per unit costs
// 2.9 ms on the C2 batch (0.8 ms with a static schedule): they are served at a
// low device-wide rate whatever the address.  So the device counter is
// claimed per workgroup, kClaim units at a time, and the workgroup's waves
// take units one by one from a 64-bit LDS word packing (end << 32 | next) of
// the chunk in hand.  One wave at a time refills (the LDS lock: 0 free,
// 1 refilling, 2 device counter exhausted); the others sleep until the new
// chunk is in, so no wave ever holds more than one unit of the tail.  Chunks
// of 8 keep the units in flight close together in the unit order (neighbours
// share cache lines of the same lists): C2 batch 0.77 ms at 8, 0.84 at 16,
// 1.21 at 64, and 0.90 at 4, where the device atomics dominate again.
// Returns the next unit of this wave, or kNone once the queue is drained.
struct ClaimLds {
    unsigned long long chunk;
    uint32_t lock;
};

__device__ uint32_t claim_unit(const ProbeArgs& A, ClaimLds* q)
{
    uint32_t u = kNone;
    if (lane_id() == 0) {
        for (;;) {
            const unsigned long long old = atomicAdd(&q->chunk, 1ull);
            if ((uint32_t)old < (uint32_t)(old >> 32)) {
                u = (uint32_t)old;
                break;
            }
            if (atomicCAS(&q->lock, 0u, 1u) == 0u) {
                // this wave refills, unless a chunk came in since its take
                const unsigned long long cur = __hip_atomic_load(&q->chunk, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
                if ((uint32_t)cur < (uint32_t)(cur >> 32)) {
                    __hip_atomic_store(&q->lock, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                    continue;
                }
                const uint32_t g = atomicAdd(A.counter, kClaim);
                if (g >= A.nunits) {
                    __hip_atomic_store(&q->lock, 2u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                    break;
                }
                const uint32_t ge = min(g + kClaim, A.nunits);
                __hip_atomic_exchange(&q->chunk, ((unsigned long long)ge << 32) | (g + 1u), __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_WORKGROUP);
                __hip_atomic_store(&q->lock, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                u = g;
                break;
            }
            uint32_t l;
            while ((l = __hip_atomic_load(&q->lock, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP)) == 1u)
                __builtin_amdgcn_s_sleep(2);
            if (l == 2u) break;
        }
    }
    return uni(__shfl(u, 0));
}

// Eight waves per SIMD (64 VGPRs; one 64-thread workgroup with 5,120 bytes
// of LDS per unit): resident waves hide the load rounds as long as the
// spilled registers stay few (C2 batch, an earlier build: 7 waves 0.88 ms,
// 8 waves 0.85 ms with 32 spilled bytes per lane; ParaPFD now spills 72
// bytes and still measured faster at 8 than at 7 or 6).  The persistent
// kernel is compiled for 7, its 72 VGPRs holding the claim loop's state.
template <int CODEC, int SEG, bool PERSIST>
__global__ __launch_bounds__(64 * kProbeWaves) __attribute__((amdgpu_waves_per_eu(PERSIST ? 7 : 8, PERSIST ? 7 : 8))) void k_probe(ProbeArgs A)
{
    // 2 KiB-aligned: every wave's window bitmap then starts at a 2 KiB boundary (bitmap_of)
    __shared__ __attribute__((aligned(2048))) uint32_t lds[kProbeWaves * wave_lds_words<CODEC>()];
    const uint32_t wid = uni(threadIdx.x >> 6);
    uint32_t* L = lds + wid * wave_lds_words<CODEC>();
    if constexpr (PERSIST) {
        __shared__ ClaimLds q;
        if (threadIdx.x == 0) {
            q.chunk = 0;
            q.lock = 0;
        }
        __syncthreads();
        for (uint32_t u; (u = claim_unit(A, &q)) != kNone;) run_unit<CODEC, SEG>(A, u, L);
    } else {
        const uint32_t u = blockIdx.x * kProbeWaves + wid;
        if (u < A.nunits) run_unit<CODEC, SEG>(A, u, L);
    }
}

// One thread per work unit: the blocks of the unit's second-shortest list
// that can hold a docID of its slice.  The slice lies in [head of its first
// block, head of the block after it - 1] of the shortest list, so two
// branchless upper-bound searches over the second list's heads bound the run
// of blocks the probe must look at, off the probe kernel's critical path.
template <int SEG>
__global__ __launch_bounds__(256) void k_unit_prologue(ProbeArgs A)
{
    const uint32_t u = blockIdx.x * 256 + threadIdx.x;
    if (u >= A.nunits) return;
    const UnitRec ur = A.units[u];
    if (ur.tc < 2) return;
    constexpr uint32_t U = kChunk / SEG;
    const TermRec R0 = A.qrecs[ur.toff], R1 = A.qrecs[ur.toff + 1];
    const uint32_t b0 = ur.k * U, nb = min(U, R0.nseg - b0);
    const uint32_t* h0 = A.ix.seg_head + R0.segnum;
    const uint32_t vmin = h0[b0];
    const uint32_t vmax = (b0 + nb < R0.nseg) ? h0[b0 + nb] - 1u : 0xffffffffu;
    const uint32_t* h1 = A.ix.seg_head + R1.segnum;
    const uint32_t n = R1.nseg;
    // upper bounds of vmin and vmax (heads <= v) are searched in [lo, lo + len]:
    // the whole list, or, for the fitted codecs, the window the list fit gives
    // once two bracket loads confirm it holds the bound (a window that misses
    // falls back to the whole list, so the fit's rounding never matters)
    uint32_t lo_a = 0, len_a = n, lo_b = 0, len_b = n;
    if ((A.ix.codec == EXC_SEGLRC || A.ix.codec == EXC_LRC || A.ix.codec == EXC_PARAPFD) && n > 64) {
        int64_t wa0, wb0, wa1, wb1;
        if (A.ix.codec != EXC_LRC) {  // SegLRC list fit, ParaPFD heads line: float intercept
            fit_window<EXC_SEGLRC, SEG>(R1, vmin, wa0, wb0);
            fit_window<EXC_SEGLRC, SEG>(R1, vmax, wa1, wb1);
        } else {
            fit_window<EXC_LRC, SEG>(R1, vmin, wa0, wb0);
            fit_window<EXC_LRC, SEG>(R1, vmax, wa1, wb1);
        }
        wa0 = min(wa0, (int64_t)n);
        wb0 = min(max(wb0, wa0), (int64_t)n);
        wa1 = min(wa1, (int64_t)n);
        wb1 = min(max(wb1, wa1), (int64_t)n);
        // the bound of v lies in [wa, wb] iff head[wa - 1] <= v (or wa = 0)
        // and head[wb] > v (or wb = n)
        const uint32_t l0 = wa0 ? h1[wa0 - 1] : 0u, r0 = wb0 < n ? h1[wb0] : 0xffffffffu;
        const uint32_t l1 = wa1 ? h1[wa1 - 1] : 0u, r1 = wb1 < n ? h1[wb1] : 0xffffffffu;
        if (l0 <= vmin && (wb0 == n || r0 > vmin)) {
            lo_a = (uint32_t)wa0;
            len_a = (uint32_t)(wb0 - wa0);
        }
        if (l1 <= vmax && (wb1 == n || r1 > vmax)) {
            lo_b = (uint32_t)wa1;
            len_b = (uint32_t)(wb1 - wa1);
        }
    }
    uint32_t a = lo_a, b = lo_b;
    const uint32_t len = max(len_a, len_b);
    for (uint32_t step = len ? 1u << (31 - __builtin_clz(len)) : 0u; step; step >>= 1) {
        if (a + step <= lo_a + len_a && h1[a + step - 1] <= vmin) a += step;
        if (b + step <= lo_b + len_b && h1[b + step - 1] <= vmax) b += step;
    }
    const uint32_t first = a ? a - 1 : 0;
    A.units[u].l1_first = first;
    A.units[u].l1_count = b ? b - first : 0;
}

__global__ void k_reset(uint32_t* c) { if (threadIdx.x == 0) *c = 0; }

__global__ __launch_bounds__(256) void k_zero_words(uint32_t* p, uint32_t n)
{
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) p[i] = 0;
}

// ---- batch planning on the device (the allocation-free entry point)

// Thread per query: the query's terms sorted by ascending list length with a
// stable insertion sort (the rule of LRC_Intersection.cu:157-179, as
// plan_queries applies it on the host), their term records copied in that
// order, and its QueryRec with the unit count, ratio class and knobs.  The
// shortest-list length is parked in short_off until k_na_scan turns the
// counts into offsets.
__global__ __launch_bounds__(256) void k_na_queries(NaPlanArgs a)
{
    const uint32_t q = blockIdx.x * 256 + threadIdx.x;
    if (q >= a.nq) return;
    const uint64_t o = a.qoff[q] - a.qoff[0];
    const uint32_t tc = (uint32_t)(a.qoff[q + 1] - a.qoff[q]);  // <= kNaMaxTerms, checked on the host
    uint32_t t[kNaMaxTerms], n[kNaMaxTerms];
    for (uint32_t i = 0; i < tc; ++i) {
        const uint32_t x = a.qterms[o + i], nx = a.terms[x].n;
        uint32_t k = i;
        while (k && nx < n[k - 1]) {
            t[k] = t[k - 1];
            n[k] = n[k - 1];
            --k;
        }
        t[k] = x;
        n[k] = nx;
    }
    for (uint32_t i = 0; i < tc; ++i) a.qrecs[o + i] = a.terms[t[i]];
    uint32_t klass = kRatioClasses - 1;
    if (tc >= 2) {
        // ratio_class of the host planner, in the same double arithmetic
        double r = (double)n[1] / (double)max(n[0], 1u);
        klass = 0;
        while (klass < 4 && r >= 10.0) {
            r /= 10.0;
            ++klass;
        }
    }
    const uint32_t n0 = tc ? n[0] : 0u;
    QueryRec Q;
    Q.toff = (uint32_t)o;
    Q.tc = tc;
    Q.first_unit = 0;
    Q.nunits = (n0 + kChunk - 1) / kChunk;
    Q.short_off = n0;
    Q.knobs = a.knobs[klass];
    Q.klass = klass;
    a.queries[q] = Q;
}

// One workgroup: exclusive sums of the queries' unit counts and shortest-list
// lengths into first_unit and short_off (each thread a contiguous range of
// queries, a block scan over the threads' totals in LDS).
__global__ __launch_bounds__(1024) void k_na_scan(NaPlanArgs a)
{
    __shared__ uint32_t su[1024];
    __shared__ uint64_t ss[1024];
    const uint32_t t = threadIdx.x, per = (a.nq + 1023) / 1024;
    const uint32_t q0 = min(a.nq, t * per), q1 = min(a.nq, q0 + per);
    uint32_t u = 0;
    uint64_t s = 0;
    for (uint32_t q = q0; q < q1; ++q) {
        u += a.queries[q].nunits;
        s += a.queries[q].short_off;
    }
    su[t] = u;
    ss[t] = s;
    __syncthreads();
    for (uint32_t d = 1; d < 1024; d <<= 1) {  // inclusive Hillis-Steele scan
        const uint32_t xu = t >= d ? su[t - d] : 0u;
        const uint64_t xs = t >= d ? ss[t - d] : 0ull;
        __syncthreads();
        su[t] += xu;
        ss[t] += xs;
        __syncthreads();
    }
    u = su[t] - u;
    s = ss[t] - s;
    for (uint32_t q = q0; q < q1; ++q) {
        QueryRec& Q = a.queries[q];
        const uint32_t nu = Q.nunits;
        const uint64_t n0 = Q.short_off;
        Q.first_unit = u;
        Q.short_off = s;
        u += nu;
        s += n0;
    }
}

// One wave per query: its work units, in query order (also the launch order).
__global__ __launch_bounds__(256) void k_na_units(NaPlanArgs a)
{
    const uint32_t q = blockIdx.x * kWaves + (threadIdx.x >> 6);
    if (q >= a.nq) return;
    const QueryRec Q = a.queries[q];
    for (uint32_t k = lane_id(); k < Q.nunits; k += 64) {
        UnitRec U;
        U.q = q;
        U.k = k;
        U.toff = Q.toff;
        U.tc = (uint16_t)Q.tc;
        U.knobs = (uint16_t)Q.knobs;
        U.l1_first = 0;
        U.l1_count = kUnset;
        U.slot = Q.first_unit + k;
        U.soff = (uint32_t)(Q.short_off + (uint64_t)k * kChunk);
        a.units[Q.first_unit + k] = U;
    }
}

// ---- scan over unit counts (replaces cudppScan + ucntResult)

// Exclusive scan in place of v[0, n) by the whole block of NT threads;
// v[n] = total.
template <int NT>
__device__ void block_excl_scan(uint32_t* v, uint32_t n)
{
    __shared__ uint32_t wsum[NT / 64];
    const uint32_t t = threadIdx.x, ln = lane_id(), wid = t >> 6;
    const uint32_t per = (n + NT - 1) / NT;
    const uint32_t b0 = t * per;
    uint32_t tot = 0;
    for (uint32_t i = 0; i < per; ++i)
        if (b0 + i < n) tot += v[b0 + i];
    const uint32_t incl = wave_incl_scan(tot);
    if (ln == 63) wsum[wid] = incl;
    __syncthreads();
    uint32_t woff = 0;
    for (uint32_t w = 0; w < wid; ++w) woff += wsum[w];
    uint32_t run = woff + incl - tot;
    __syncthreads();
    for (uint32_t i = 0; i < per; ++i) {
        if (b0 + i < n) {
            const uint32_t x = v[b0 + i];
            v[b0 + i] = run;
            run += x;
        }
    }
    if (t == NT - 1) v[n] = woff + incl;
}

// Each block scans kScanBlockItems unit counts and publishes its sum, then
// takes a ticket (device-scope fence first, so the sum is visible before the
// ticket is).  The block holding the last ticket turns the block sums into
// exclusive offsets in place (bsum[nblk] = total) and resets the ticket for
// the next launch.  No block waits for another; the ticket is zeroed when the
// plan is uploaded and every launch leaves it at zero.
__global__ __launch_bounds__(256) void k_scan_local(ScanArgs a)
{
    __shared__ uint32_t wsum[4];
    __shared__ uint32_t last;
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
    if (threadIdx.x == 255) {
        a.bsum[blockIdx.x] = woff + incl;
        __threadfence();
        last = atomicAdd(a.ticket, 1u) == gridDim.x - 1 ? 1u : 0u;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();  // every block's sum is visible past the last ticket
    block_excl_scan<256>(a.bsum, gridDim.x);
    if (threadIdx.x == 0) *a.ticket = 0;
}

// bsum[0] = 0 for a batch without units
__global__ __launch_bounds__(64) void k_scan_empty(uint32_t* bsum)
{
    if (threadIdx.x == 0) bsum[0] = 0;
}

__device__ __forceinline__ uint32_t unit_excl(const ScanArgs& a, uint32_t u)
{
    return u < a.nunits ? a.uoff[u] + a.bsum[u / kScanBlockItems] : a.bsum[(a.nunits + kScanBlockItems - 1) / kScanBlockItems];
}

// One wave per kScatterUnits consecutive units (query order).  Their outputs
// form one contiguous range, the scan's offsets being exclusive sums: the wave
// loads the units' counts and stage slots and the first unit's offset in one
// load round, then copies the range in coalesced rows, each lane taking its
// element from the last unit whose offset inside the range is not above it,
// with kScatterRows rows read before any is written.  (One wave per unit was
// two dependent load rounds for ~37 survivors on average.)
__global__ __launch_bounds__(256) void k_scatter(ScanArgs a)
{
    const uint32_t ln = lane_id();
    const uint32_t u0 = (blockIdx.x * kWaves + (threadIdx.x >> 6)) * kScatterUnits;
    if (u0 < a.nunits) {
        // loads are unconditional (clamped indexes, results selected after):
        // exec-masked loads make the compiler wait for every outstanding
        // access before each later store
        const uint32_t base = uni(a.uoff[u0] + a.bsum[u0 / kScanBlockItems]);
        const uint32_t u = u0 + ln;
        const bool mine = ln < kScatterUnits && u < a.nunits;
        const uint32_t uc = min(u0 + min(ln, kScatterUnits - 1), a.nunits - 1);
        const uint32_t c_raw = a.ucount[uc], s_raw = a.units[uc].soff;
        const uint32_t cnt = mine ? c_raw : 0u;
        const uint32_t soff = mine ? s_raw : 0u;
        // each unit's offset inside the range and its stage slot minus that
        // offset (modulo 2^32: the sum with an element index is the slot)
        uint32_t ex[kScatterUnits], sh[kScatterUnits], total = 0;
#pragma unroll
        for (int j = 0; j < (int)kScatterUnits; ++j) {
            ex[j] = total;
            sh[j] = uni(__builtin_amdgcn_readlane(soff, j)) - total;
            total += uni(__builtin_amdgcn_readlane(cnt, j));
        }
        const uint32_t* __restrict__ src = a.stage;
        uint32_t* __restrict__ dst = a.result + base;
        for (uint32_t r0 = 0; r0 < total; r0 += 64u * kScatterRows) {
            uint32_t x[kScatterRows];
#pragma unroll
            for (int i = 0; i < (int)kScatterRows; ++i) {
                const uint32_t e = r0 + ln + 64u * (uint32_t)i;
                uint32_t d = sh[0];
#pragma unroll
                for (int j = 1; j < (int)kScatterUnits; ++j) d = e >= ex[j] ? sh[j] : d;
                x[i] = src[e < total ? e + d : 0u];
            }
#pragma unroll
            for (int i = 0; i < (int)kScatterRows; ++i) {
                const uint32_t e = r0 + ln + 64u * (uint32_t)i;
                if (e < total) dst[e] = x[i];
            }
        }
    }
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < a.nq) {
        const QueryRec Q = a.queries[g];
        a.qcum[g] = unit_excl(a, Q.first_unit + Q.nunits);
    }
}

// ---- whole-list decompression of the ParaPFor-layout codecs (ParaPFD, SegLRC, LRC)
// One wave per unit of kDecodeUnit docIDs.  The unit's block records come in
// one lane-parallel load (lane nb: the next block's, whose offset ends the
// run; the list's last unit ends at the list's word count instead); the packed
// words of all its blocks -- one contiguous run of the list -- then come in one
// round of 16-byte loads into the wave's LDS.  LRC / SegLRC: every block's
// fields go from LDS to registers and each value is formed and stored
// straight to HBM (C2 SegLRC-128 0.561 -> 0.405 ms).  ParaPFD: all the unit's
// blocks advance phase by phase (fields and exceptions from LDS to registers,
// stage, patch, read back, segHead + scan, values back through LDS, coalesced
// rows out), so their LDS round trips overlap instead of forming one chain per
// block: C2 ParaPFD-128 0.826 -> 0.57 ms, ParaPFD-256 0.722 -> 0.42 ms.  A run
// longer than kGapWords words (or a list whose word count is unknown, or a
// ParaPFD block with more than 64 exceptions) decodes its blocks from HBM one
// at a time, staged in the same buffer.  (NewPFD keeps k_decompress: the same
// phase ordering with the field-parallel Simple16 unpack of two blocks'
// streams measured 1.17 ms against 0.83.)
constexpr uint32_t kGapWords = 1024;
constexpr uint32_t kGapWaves = 4;  // waves per workgroup of the run decoder
// waves per SIMD each instance is compiled for (its phase path's registers
// without spills; measured on the C2 lists)
template <int CODEC, int SEG>
constexpr int kGapOcc = !gap_codec(CODEC) ? 8 : SEG == 256 ? 7 : 5;

template <int CODEC, int SEG>
__global__ __launch_bounds__(64 * kGapWaves) __attribute__((amdgpu_waves_per_eu(kGapOcc<CODEC, SEG>, kGapOcc<CODEC, SEG>))) void k_decode_run(DevIndex ix, const DecUnit* units, uint32_t nunits, uint32_t* out)
{
    static_assert(CODEC == EXC_PARAPFD || CODEC == EXC_SEGLRC || CODEC == EXC_LRC, "block codecs with a ParaPFor layout");
    constexpr uint32_t U = kDecodeUnit / SEG;
    constexpr int V = SEG / 64;
    // + 128: a block's fields are read for every position below SEG (up to
    // 124 words past a short last block)
    static_assert(SEG <= kGapWords, "the per-block path stages in the run's buffer");
    __shared__ __attribute__((aligned(16))) uint32_t lds_buf[kGapWaves][kGapWords + 128];
    const uint32_t wid = uni(threadIdx.x >> 6), u = blockIdx.x * kGapWaves + wid, ln = lane_id();
    uint32_t* const buf = lds_buf[wid];
    if (u >= nunits) return;
    const DecUnit D = units[u];
    const uint32_t k = uni(D.k), n = uni(D.n);
    const uint32_t nseg = (n + SEG - 1) / SEG;
    const uint32_t nb = uni(min(U, nseg - k));
    const bool has_next = k + nb < nseg;
    TermRec R;  // the fields seg_len, the LRC fit and the per-block path read
    R.woff = D.woff;
    R.n = n;
    R.segnum = D.segnum;
    R.nseg = nseg;
    R.slope = D.slope;
    R.icpt = D.icpt;
    const SegRec gr = ix.rec[D.segnum + k + min(ln, has_next ? nb : nb - 1u)];
    const uint32_t w_first = uni(__builtin_amdgcn_readlane(gr.off, 0));
    const uint32_t w_end = has_next ? uni(__builtin_amdgcn_readlane(gr.off, (int)nb)) : uni(D.wend);
    uint32_t* o = out + D.out;
    // the run is copied in 16-byte pieces from the 16-byte boundary at or
    // before its first word; that word then sits at buf[sh]
    const uint64_t gfirst = D.woff + w_first;
    const uint32_t sh = (uint32_t)(gfirst & 3u);
    const bool lds_run = w_end != kUnset && w_end >= w_first && w_end - w_first + sh <= kGapWords &&
                         ((uintptr_t)ix.words & 15u) == 0;
    if (lds_run) {
        constexpr uint32_t kRounds = kGapWords / 256 + 1;
        const uint32_t nq = (w_end - w_first + sh + 3) >> 2;
        const uint4* src = reinterpret_cast<const uint4*>(ix.words + (gfirst - sh));
        uint4 x[kRounds];
#pragma unroll
        for (uint32_t i = 0; i < kRounds; ++i) {
            const uint32_t q = ln + 64 * i;
            x[i] = q < nq ? src[q] : make_uint4(0, 0, 0, 0);
        }
#pragma unroll
        for (uint32_t i = 0; i < kRounds; ++i) {
            const uint32_t q = ln + 64 * i;
            if (q < nq) reinterpret_cast<uint4*>(buf)[q] = x[i];
        }
        wave_sync();
    }
    // The unit's blocks advance phase by phase, all of them in each phase, so
    // their LDS reads, patches and scans overlap instead of forming one chain
    // per block.  Fields and each block's exceptions (at most 64) go to
    // registers first; the run's LDS copy is then dead and its words become
    // the unit's output staging.  Units with more exceptions in a block take
    // the per-block path below.
    // LRC / SegLRC: every block's fields from the LDS copy to registers, then
    // each value formed and stored straight to HBM (lane l holds positions l,
    // l + 64, ...: coalesced rows)
    if constexpr (!gap_codec(CODEC)) {
        if (lds_run) {
            uint32_t f[U][V];
            SegRec rec[U];
#pragma unroll
            for (uint32_t g = 0; g < U; ++g) {
                const int gs = (int)min(g, nb - 1u);  // past nb: a copy of the last block, not used
                rec[g].off = uni(__builtin_amdgcn_readlane(gr.off, gs));
                rec[g].desc = uni(__builtin_amdgcn_readlane(gr.desc, gs));
                rec[g].slope = __uint_as_float(uni(__builtin_amdgcn_readlane(__float_as_uint(gr.slope), gs)));
                rec[g].icpt = __uint_as_float(uni(__builtin_amdgcn_readlane(__float_as_uint(gr.icpt), gs)));
                rec[g].base = (int32_t)uni(__builtin_amdgcn_readlane((uint32_t)gr.base, gs));
                const uint32_t* data = buf + sh + (rec[g].off - w_first) + 1;
                load_fields<CODEC, SEG, false>(data, desc_lb(rec[g].desc), SEG, f[g]);
            }
#pragma unroll
            for (uint32_t g = 0; g < U; ++g) {
                if (g < nb) {
                    const uint32_t s = k + g, len = seg_len(R, s, SEG);
                    const float icv = vgpr_copy(rec[g].icpt);
#pragma unroll
                    for (int t = 0; t < V; ++t) {
                        const uint32_t j = ln + 64 * t;
                        uint32_t x;
                        if constexpr (CODEC == EXC_SEGLRC)
                            x = f[g][t] + seg_pred(j, rec[g].slope, icv, rec[g].base);
                        else
                            x = f[g][t] + lrc_pred(s * SEG + j, R.slope, R.icpt);
                        if (j < len) o[g * SEG + j] = x;
                    }
                }
            }
            return;
        }
    }
    const bool few_exc = !__ballot(ln < nb && desc_en(gr.desc) > 64u);
    if constexpr (gap_codec(CODEC) && SEG * U <= kGapWords) {
        if (lds_run && few_exc) {
            uint32_t f[U][V], xq[U], xh[U], lbv[U], lenv[U];
#pragma unroll
            for (uint32_t g = 0; g < U; ++g) {
                const uint32_t gs = min(g, nb - 1u);  // past nb: a copy of the last block, not used
                const uint32_t off = uni(__builtin_amdgcn_readlane(gr.off, (int)gs));
                const uint32_t desc = uni(__builtin_amdgcn_readlane(gr.desc, (int)gs));
                const uint32_t lb = desc_lb(desc), en = desc_en(desc);
                const uint32_t len = seg_len(R, k + gs, SEG);
                const uint32_t* data = buf + sh + (off - w_first) + 1;
                lbv[g] = lb;
                lenv[g] = len;
#pragma unroll
                for (int t = 0; t < V; ++t) f[g][t] = bits_at(data, lb, ln * V + t);
                const uint32_t* pp = data + words_for(len, lb);
                const uint32_t eb = desc_eb(desc), hb = desc_hb(desc), e = min(ln, en ? en - 1u : 0u);
                xq[g] = ln < en ? bits_at(pp, eb, e) : kNone;
                xh[g] = bits_at(pp + words_for(en, eb), hb, e);
            }
            {
                wave_sync();
                uint32_t* L = buf;  // output staging: block g at L + g * SEG
#pragma unroll
                for (uint32_t g = 0; g < U; ++g)
#pragma unroll
                    for (int t = 0; t < V; ++t) L[g * SEG + ln * V + t] = f[g][t];
                wave_sync();
                // exception e of block g: field[pos_e] |= high_e << lb (positions are distinct)
#pragma unroll
                for (uint32_t g = 0; g < U; ++g)
                    if (g < nb && xq[g] < lenv[g])
                        __hip_atomic_fetch_or(L + g * SEG + xq[g], xh[g] << lbv[g], __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_WAVEFRONT);
                wave_sync();
#pragma unroll
                for (uint32_t g = 0; g < U; ++g) {
                    const uint32_t head = uni(__builtin_amdgcn_readlane(gr.head, (int)min(g, nb - 1u)));
                    uint32_t acc = 0;
#pragma unroll
                    for (int t = 0; t < V; ++t) {
                        uint32_t x = L[g * SEG + ln * V + t];
                        if (ln == 0 && t == 0) x = head;
                        acc += x;
                        f[g][t] = acc;
                    }
                    const uint32_t excl = wave_incl_scan(acc) - acc;
#pragma unroll
                    for (int t = 0; t < V; ++t) f[g][t] += excl;
                }
                wave_sync();
#pragma unroll
                for (uint32_t g = 0; g < U; ++g)
#pragma unroll
                    for (int t = 0; t < V; ++t) L[g * SEG + ln * V + t] = f[g][t];
                wave_sync();
                const uint32_t total = (nb - 1u) * SEG + seg_len(R, k + nb - 1u, SEG);
                for (uint32_t e = ln; e < total; e += 64) o[e] = L[e];
                return;
            }
        }
    }
    for (uint32_t g = 0; g < nb; ++g) {
        SegRec rec;
        rec.off = uni(__builtin_amdgcn_readlane(gr.off, (int)g));
        rec.desc = uni(__builtin_amdgcn_readlane(gr.desc, (int)g));
        rec.head = uni(__builtin_amdgcn_readlane(gr.head, (int)g));
        rec.slope = __uint_as_float(uni(__builtin_amdgcn_readlane(__float_as_uint(gr.slope), (int)g)));
        rec.icpt = __uint_as_float(uni(__builtin_amdgcn_readlane(__float_as_uint(gr.icpt), (int)g)));
        rec.rl = rec.rr = 0;
        rec.base = (int32_t)uni(__builtin_amdgcn_readlane((uint32_t)gr.base, (int)g));
        const uint32_t s = k + g, len = seg_len(R, s, SEG);
        // the block from HBM, staged in the buffer
        const uint32_t* data = ix.words + R.woff + rec.off + 1;
        uint32_t* const stage = buf;
        uint32_t f[V];
        load_fields<CODEC, SEG>(data, desc_lb(rec.desc), len, f);
        wave_sync();  // the previous block's (or the phase path's) reads of the buffer come first
        finish_block<CODEC, SEG>(R, s, rec, len, data, f, stage, nullptr);
        wave_sync();
        for (uint32_t e = ln; e < len; e += 64) o[g * SEG + e] = stage[e];
        wave_sync();
    }
}

// ---- whole-list decompression (the *_Decompression programs)
// One wave per unit of kDecodeUnit docIDs: a run of kDecodeUnit/SEG blocks of
// one list (the HS codec: one chunk of SEG elements).  Every block record and
// packed word of the run is requested before any value is formed; LRC and
// SegLRC values go from registers straight to HBM in coalesced stores,
// ParaPFD and NewPFD values pass through LDS for the exception patch and the
// scan (values stored straight from registers, V consecutive per lane, and a
// single patch area measured slower: 0.99 vs 0.91 ms for the C2 ParaPFD index).
template <int CODEC, int SEG>
__global__ __launch_bounds__(256) void k_decompress(DevIndex ix, const DecUnit* units, uint32_t nunits, uint32_t* out)
{
    constexpr bool HS = CODEC == EXC_HS_SEGLRC;
    constexpr int U = HS ? 1 : (int)(kDecodeUnit / SEG);
    constexpr bool VIA_LDS = HS || gap_codec(CODEC);
    constexpr int W = U * SEG + (CODEC == EXC_NEWPFD ? 2 * SEG : 0);  // words per wave
    __shared__ uint32_t lds[VIA_LDS ? kWaves * W : 1];
    const uint32_t wid = uni(threadIdx.x >> 6), ln = lane_id();
    const uint32_t u = blockIdx.x * kWaves + wid;
    if (u >= nunits) return;
    const DecUnit D = units[u];
    const uint32_t k = uni(D.k);
    uint32_t* o = out + D.out;
    if constexpr (raw_codec(CODEC)) {
        const uint32_t b0 = k * SEG, n = min((uint32_t)(U * SEG), D.n - b0);
        const uint32_t* src = ix.words + D.woff + b0;
        for (uint32_t e = ln; e < n; e += 64) o[e] = src[e];
    } else if constexpr (HS) {
        const TermRec R = ix.terms[uni(D.term)];
        uint32_t* L = lds + wid * SEG;
        const uint32_t n = decode_hs_chunk<SEG>(ix, R, k, L);
        wave_sync();
        for (uint32_t e = ln; e < n; e += 64) o[e] = L[e];
    } else {
        TermRec R;  // the fields decode_blocks reads, from the unit record
        R.woff = D.woff;
        R.n = D.n;
        R.segnum = D.segnum;
        R.nseg = (D.n + SEG - 1) / SEG;
        R.slope = D.slope;
        R.icpt = D.icpt;
        const uint32_t nb = uni(min((uint32_t)U, R.nseg - k));
        const auto sid = [&](int g) { return k + (uint32_t)g; };
        if constexpr (VIA_LDS) {
            uint32_t* L = lds + wid * W;
            decode_blocks<CODEC, SEG, U, true, true>(ix, R, nb, sid, L);
            wave_sync();
            const uint32_t n = (nb - 1) * SEG + seg_len(R, k + nb - 1, SEG);
            for (uint32_t e = ln; e < n; e += 64) o[e] = L[e];
        } else {
            decode_blocks<CODEC, SEG, U, true>(ix, R, nb, sid, o);
        }
    }
}

template <int CODEC>
hipError_t probe_for_seg(const ProbeArgs& a, uint32_t grid, hipStream_t s)
{
    if (a.persistent) {
        if (a.ix.seg == 256)
            hipLaunchKernelGGL((k_probe<CODEC, 256, true>), dim3(grid), dim3(64 * kProbeWaves), 0, s, a);
        else
            hipLaunchKernelGGL((k_probe<CODEC, 128, true>), dim3(grid), dim3(64 * kProbeWaves), 0, s, a);
    } else {
        if (a.ix.seg == 256)
            hipLaunchKernelGGL((k_probe<CODEC, 256, false>), dim3(grid), dim3(64 * kProbeWaves), 0, s, a);
        else
            hipLaunchKernelGGL((k_probe<CODEC, 128, false>), dim3(grid), dim3(64 * kProbeWaves), 0, s, a);
    }
    return hipGetLastError();
}

template <int CODEC>
hipError_t decomp_for_seg(const DevIndex& ix, const DecUnit* units, uint32_t n, uint32_t* out, hipStream_t s)
{
    if constexpr (CODEC == EXC_PARAPFD || CODEC == EXC_SEGLRC || CODEC == EXC_LRC) {
        const uint32_t grid = (n + kGapWaves - 1) / kGapWaves;
        if (ix.seg == 256)
            hipLaunchKernelGGL((k_decode_run<CODEC, 256>), dim3(grid), dim3(64 * kGapWaves), 0, s, ix, units, n, out);
        else
            hipLaunchKernelGGL((k_decode_run<CODEC, 128>), dim3(grid), dim3(64 * kGapWaves), 0, s, ix, units, n, out);
        return hipGetLastError();
    }
    const uint32_t grid = (n + kWaves - 1) / kWaves;
    if (ix.seg == 256)
        hipLaunchKernelGGL((k_decompress<CODEC, 256>), dim3(grid), dim3(256), 0, s, ix, units, n, out);
    else
        hipLaunchKernelGGL((k_decompress<CODEC, 128>), dim3(grid), dim3(256), 0, s, ix, units, n, out);
    return hipGetLastError();
}

template <int CODEC>
int occ_for(uint32_t seg)
{
    int nb = 0;
    hipError_t e = (seg == 256)
        ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(&k_probe<CODEC, 256, true>), 64 * kProbeWaves, 0)
        : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(&k_probe<CODEC, 128, true>), 64 * kProbeWaves, 0);
    return (e == hipSuccess && nb > 0) ? nb : 4;
}

}  // namespace

uint32_t probe_waves_per_block() { return kProbeWaves; }

int probe_blocks_per_cu(int codec, uint32_t seg)
{
    switch (codec) {
    case EXC_LRC: return occ_for<EXC_LRC>(seg);
    case EXC_SEGLRC: return occ_for<EXC_SEGLRC>(seg);
    case EXC_HS_SEGLRC: return occ_for<EXC_HS_SEGLRC>(seg);
    case EXC_PARAPFD: return occ_for<EXC_PARAPFD>(seg);
    case EXC_RAW: return occ_for<EXC_RAW>(seg);
    case EXC_RAW_HS: return occ_for<EXC_RAW_HS>(seg);
    }
    return 4;
}

hipError_t launch_reset(uint32_t* counter, hipStream_t s)
{
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, counter);
    return hipGetLastError();
}

hipError_t launch_zero_words(uint32_t* p, uint32_t n, hipStream_t s)
{
    if (!n) return hipSuccess;
    const uint32_t grid = min((n + 255) / 256, 1024u);
    hipLaunchKernelGGL(k_zero_words, dim3(grid), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_noalloc_plan(const NaPlanArgs& a, hipStream_t s)
{
    if (!a.nq) return hipSuccess;
    hipLaunchKernelGGL(k_na_queries, dim3((a.nq + 255) / 256), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_na_scan, dim3(1), dim3(1024), 0, s, a);
    hipLaunchKernelGGL(k_na_units, dim3((a.nq + kWaves - 1) / kWaves), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_probe(const ProbeArgs& a, uint32_t grid, hipStream_t s)
{
    switch (a.ix.codec) {
    case EXC_LRC: return probe_for_seg<EXC_LRC>(a, grid, s);
    case EXC_SEGLRC: return probe_for_seg<EXC_SEGLRC>(a, grid, s);
    case EXC_HS_SEGLRC: return probe_for_seg<EXC_HS_SEGLRC>(a, grid, s);
    case EXC_PARAPFD: return probe_for_seg<EXC_PARAPFD>(a, grid, s);
    case EXC_RAW: return probe_for_seg<EXC_RAW>(a, grid, s);
    case EXC_RAW_HS: return probe_for_seg<EXC_RAW_HS>(a, grid, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_unit_prologue(const ProbeArgs& a, hipStream_t s)
{
    // HS probes find buckets directly; raw lists have no block heads
    if (!a.nunits || a.ix.codec == EXC_HS_SEGLRC || raw_codec(a.ix.codec)) return hipSuccess;
    const uint32_t grid = (a.nunits + 255) / 256;
    if (a.ix.seg == 256)
        hipLaunchKernelGGL(k_unit_prologue<256>, dim3(grid), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(k_unit_prologue<128>, dim3(grid), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_scan(const ScanArgs& a, hipStream_t s)
{
    const uint32_t nblk = (a.nunits + kScanBlockItems - 1) / kScanBlockItems;
    if (nblk)
        hipLaunchKernelGGL(k_scan_local, dim3(nblk), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(k_scan_empty, dim3(1), dim3(64), 0, s, a.bsum);
    const uint32_t per_block = kWaves * kScatterUnits;
    const uint32_t g1 = (a.nunits + per_block - 1) / per_block, g2 = (a.nq + 255) / 256;
    const uint32_t grid = g1 > g2 ? g1 : g2;
    if (grid) hipLaunchKernelGGL(k_scatter, dim3(grid), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_decompress(const DevIndex& ix, const DecUnit* units, uint32_t n, uint32_t* out, hipStream_t s)
{
    if (!n) return hipSuccess;
    switch (ix.codec) {
    case EXC_LRC: return decomp_for_seg<EXC_LRC>(ix, units, n, out, s);
    case EXC_SEGLRC: return decomp_for_seg<EXC_SEGLRC>(ix, units, n, out, s);
    case EXC_HS_SEGLRC: return decomp_for_seg<EXC_HS_SEGLRC>(ix, units, n, out, s);
    case EXC_PARAPFD: return decomp_for_seg<EXC_PARAPFD>(ix, units, n, out, s);
    case EXC_RAW: return decomp_for_seg<EXC_RAW>(ix, units, n, out, s);
    case EXC_RAW_HS: return decomp_for_seg<EXC_RAW_HS>(ix, units, n, out, s);
    case EXC_NEWPFD: return decomp_for_seg<EXC_NEWPFD>(ix, units, n, out, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace exc
