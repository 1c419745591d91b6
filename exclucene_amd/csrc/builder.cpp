// builder.cpp -- CPU index builder and synthetic corpus generator.
//
// Replaces the reference's two offline programs per codec: the *_Generator
// (double-precision fits + residual kernel + min shift) and the *_Compression
// (ParaPFor packing + side files).  Output arrays follow the reference's file
// layouts exactly (GC = src/gpu/Compressed/src, GI = src/gpu/Compressed/include):
//   fits        GC/LRC/LRC_Generator/LRC_Generator.cu:221-287,
//               GC/SegLRC/SegLRC_Generator/SegLRC_Generator.cu:264-331
//   residuals   GC/LRC/LRC_Generator/LRC_Generator_kernel.cu:55-91,
//               GC/SegLRC/SegLRC_Generator/SegLRC_Generator_kernel.cu:62-100
//   min shift   GC/LRC/LRC_Generator/LRC_Generator.cu:165-180,
//               GC/SegLRC/SegLRC_Generator/SegLRC_Generator.cu:185-233
//   HS table    GC/HS_SegLRC/HS_SegLRC_Generator/HS_SegLRC_Generator.cu:503-594
//   packing     GC/LRC/LRC_Compression/main.cpp:163-199, GI/ParaPFor.cpp:56-94,
//               GC/ParaPFD/ParaPFD_Compression/main.cpp:151-189,
//               GI/NewPFor.cpp:54-99, GI/Simple16.h:13-48,
//               GC/NewPFD/NewPFD_Compression/main.cpp:151-189,
//               GC/HS_SegLRC/HS_SegLRC_Compression/main.cpp:149-178
// Lists are encoded in parallel (static partition over std::thread workers);
// the result is identical for any thread count.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/exclucene_amd.h"
#include "exc_format.h"
#include "exc_internal.h"
#include "exc_encode.h"

using namespace exc;

static_assert(sizeof(exc_gu_regression_info) == 36, "at_regression_info rows are 36 bytes");

struct exc_built {
    int32_t codec = 0;
    uint32_t seg = 256;
    std::vector<exc_term_ind1> ind1;
    std::vector<uint32_t> ind2;
    std::vector<uint32_t> seg_num, seg_offset, seg_head, seg_median;
    std::vector<exc_lrc_info> lrc;
    std::vector<exc_seg_info> seg_info, list_info;
    std::vector<int32_t> base;
    std::vector<uint32_t> hash_offset, hash_info;
    std::vector<exc_gu_regression_info> gu_reg;
};

namespace {

// ---------------------------------------------------------------- ParaPFor

// Smallest b with #(v >= 2^b) <= frac * n (the shared rule of exc_encode.h).
uint32_t pfor_best_b(const uint32_t* v, uint32_t n, float frac)
{
    return exc::pfor_best_b([v](uint32_t i) { return v[i]; }, n, frac);
}

void put_bits(uint32_t* out, const uint32_t* v, uint32_t n, uint32_t b)
{
    uint32_t nw = words_for(n, b);
    std::memset(out, 0, (size_t)nw * 4);
    if (!b) return;
    for (uint32_t i = 0; i < n; ++i) {
        uint64_t bit = (uint64_t)i * b;
        uint32_t w = (uint32_t)(bit >> 5), s = (uint32_t)(bit & 31);
        out[w] |= v[i] << s;
        if (s + b > 32) out[w + 1] |= v[i] >> (32 - s);
    }
}

// One ParaPFor block appended to out, encoded by the code the device builder
// runs too (exc_encode.h); false if the descriptor cannot hold it.
bool pfor_block(const uint32_t* v, uint32_t n, float frac, std::vector<uint32_t>& out)
{
    const auto val = [v](uint32_t i) { return v[i]; };
    const PforPlan p = pfor_plan(val, n, frac);
    if (!p.words) return false;
    const size_t at = out.size();
    out.resize(at + p.words);
    pfor_write(val, n, p.desc, out.data() + at);
    return true;
}

// ---------------------------------------------------------------- NewPFor

// Simple16 words for vals[0..n) appended to out: at each position the first
// selector (in order 0..15) whose fields hold the next min(count, left)
// values (GI/Simple16.h:13-48).  False if a value needs more than 28 bits,
// which no selector holds (the reference's encoder never returns then).
bool s16_encode(const uint32_t* vals, uint32_t n, std::vector<uint32_t>& out)
{
    for (uint32_t i = 0; i < n;) {
        bool done = false;
        for (uint32_t k = 0; k < 16 && !done; ++k) {
            const uint32_t L = s16_layout(k);
            uint32_t width[28], cnt = 0;
            for (uint32_t r = 0; r < 3; ++r)
                for (uint32_t q = 0; q < ((L >> (10 * r)) & 31u); ++q) width[cnt++] = (L >> (10 * r + 5)) & 31u;
            const uint32_t vol = std::min(cnt, n - i);
            uint32_t word = k << 28, sh = 0, j = 0;
            for (; j < vol && (vals[i + j] >> width[j]) == 0; sh += width[j], ++j) word |= vals[i + j] << sh;
            if (j == vol) {
                out.push_back(word);
                i += vol;
                done = true;
            }
        }
        if (!done) return false;
    }
    return true;
}

// One NewPFor block appended to out (GI/NewPFor.cpp:54-99): b by the ParaPFor
// rule, low parts packed, exceptions as one Simple16 stream of position
// deltas then high parts.  When a high part would need more than 28 bits b
// grows until none does (the reference's encoder would not terminate).
bool npfd_block(const uint32_t* v, uint32_t n, float frac, std::vector<uint32_t>& out)
{
    uint32_t b = pfor_best_b(v, n, frac), mx = 0;
    for (uint32_t i = 0; i < n; ++i) mx = std::max(mx, v[i]);
    b = std::max(b, bits_of(mx) > 28 ? bits_of(mx) - 28 : 0u);
    if (b > 31) return false;
    std::vector<uint32_t> low(v, v + n), all;
    std::vector<uint32_t> pos, hi;
    for (uint32_t i = 0; i < n; ++i) {
        if (bits_of(v[i]) > b) {
            pos.push_back(i);
            hi.push_back(v[i] >> b);
            low[i] = v[i] & ((1u << b) - 1u);
        }
    }
    const uint32_t en = (uint32_t)pos.size();
    if (en > 65535) return false;
    const size_t at = out.size();
    out.resize(at + 1 + words_for(n, b));
    out[at] = (b << 26) | en;
    put_bits(out.data() + at + 1, low.data(), n, b);
    if (en) {
        all.resize(2 * (size_t)en);
        for (uint32_t i = 0; i < en; ++i) {
            all[i] = i ? pos[i] - pos[i - 1] - 1 : pos[0];
            all[en + i] = hi[i];
        }
        if (!s16_encode(all.data(), 2 * en, out)) return false;
    }
    return true;
}

// ---------------------------------------------------------------- fits

// residuals against a segment fit, shifted by their minimum; returns the base
int32_t seg_residuals(const uint32_t* L, uint32_t n, const exc_seg_info& f, uint32_t* vd)
{
    int32_t mn = 0;
    for (uint32_t j = 0; j < n; ++j) {
        int32_t d = seg_residual(L, j, f);
        vd[j] = (uint32_t)d;
        if (j == 0 || d < mn) mn = d;
    }
    for (uint32_t j = 0; j < n; ++j) vd[j] = (uint32_t)((int32_t)vd[j] - mn);
    return n ? mn : 0;
}

void hs_table(const uint32_t* L, uint32_t n, uint32_t target, uint32_t fixed_shift, std::vector<uint32_t>& t)
{
    uint32_t maxb = n ? bits_of(L[n - 1]) : 0;
    if (!fixed_shift && n <= target) { t = {maxb, 1u, 0u, n}; return; }
    uint32_t shift = fixed_shift ? fixed_shift : maxb - bits_of((n + target - 1) / target);
    uint32_t nb = n ? (L[n - 1] >> shift) + 1 : 1;
    t.assign(nb + 3, 0);
    t[0] = shift;
    t[1] = nb;
    uint32_t cur = 0;
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t b = L[i] >> shift;
        while (cur < b) t[2 + (++cur)] = i;
    }
    t[2 + nb] = n;
}

// ceil(log2 x) as the uncompressed hash-segment generator computes it, in
// double logarithms (cuda_hashSegment.cu:872-881; for x = 2^29 and 2^31 the
// quotient rounds above the integer and the result is one more).  x = 0, whose
// logarithm the reference takes anyway, gives 0 here.
uint32_t gu_ceil_log2(double x)
{
    if (!(x > 0.0)) return 0;
    const double d = std::log(x) / std::log(2.0);
    if (d <= 0.0) return 0;
    const uint32_t i = (uint32_t)d;
    return (d - (double)i > 0.0) ? i + 1 : i;
}

// The uncompressed hash-segment table of one list, as RunHashSegInfoGenerator
// writes it (cuda_hashSegment.cu:815-947): lists of at most `target` docIDs
// get [bits, 0, n] (one bucket, bits = ceil(log2 max)); longer lists get
// [shift][start_0 .. start_nb] with shift = bits - ceil(log2(n / target)),
// nb = (max >> shift) + 1 and start_b the first index whose docID >> shift
// is >= b.  A fixed shift uses the second form for every list.  An empty list
// (whose last docID the reference reads anyway) gets [shift, 0, 0].
void gu_hash_table(const uint32_t* L, uint32_t n, uint32_t target, uint32_t fixed_shift, std::vector<uint32_t>& t)
{
    if (n == 0) { t = {fixed_shift, 0u, 0u}; return; }
    const uint32_t mx = L[n - 1];
    uint32_t shift = fixed_shift;
    if (!fixed_shift) {
        const uint32_t bits = gu_ceil_log2((double)mx);
        if (n <= target) { t = {bits, 0u, n}; return; }
        shift = bits - gu_ceil_log2((double)n / target);
    }
    const uint32_t nb = (mx >> shift) + 1;
    t.assign(nb + 2, 0);
    t[0] = shift;
    uint32_t cur = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t b = L[i] >> shift;
        while (cur < b) t[1 + (++cur)] = i;
    }
    t[1 + nb] = n;
}

// The regression row of one uncompressed list, as getRegressionInfo and
// RunGetRegressionInfo compute it (GU/cuda_runRegressionAlgorithm/
// cuda_linearRegression.cu:792-873, 900-925; x magnification 1): a
// one-element list gets slope 1 and its docID as intercept, every other
// list (the empty one included, whose line is NaN) the least-squares line
// with means accumulated term by term in double, slope and intercept stored
// as float, and the ranges from each element's float position on the line.
exc_gu_regression_info gu_regression(const uint32_t* L, uint32_t n)
{
    exc_gu_regression_info r;
    std::memset(&r, 0, sizeof r);
    if (n == 1) {
        r.fSlope = 1.0f;
        r.fIntercept = (float)L[0];
        return r;
    }
    const uint32_t mag = 1;
    double xa = 0, ya = 0, sx = 0, sxy = 0, sy = 0;
    for (uint32_t i = 0; i < n; ++i) {
        xa += (double)(i * mag) / (double)n;
        ya += (double)L[i] / (double)n;
    }
    for (uint32_t i = 0; i < n; ++i) {
        const double y = (double)L[i], dx = (double)i - xa;
        sx += dx * dx;
        sxy += dx * (y - ya);
        sy += (y - ya) * (y - ya);
    }
    r.fSlope = (float)(sxy / sx);
    r.fIntercept = (float)(ya - (double)r.fSlope * xa);
    r.fRSquare = n != 2 ? (float)(sxy / std::sqrt(sx * sy)) : 1.0f;
    r.fMultiple = std::sqrt(1.0f + r.fSlope * r.fSlope);
    r.nMagnification = mag;
    float left = 0.0f, right = 0.0f;
    for (uint32_t i = 0; i < n; ++i) {
        const float x = ((float)L[i] - r.fIntercept) / r.fSlope / (float)mag;
        const float d = x - (float)i;
        if (d > left) left = d;
        else if (d < right) right = d;
    }
    r.nRangeLeft = (uint32_t)left + 1;
    r.nRangeRight = (uint32_t)(0.0f - right) + 1;
    return r;
}

// Encoded form of one list before it is placed in the index.
struct ListOut {
    std::vector<uint32_t> words, seg_offset, seg_head, seg_median, hash;
    std::vector<exc_seg_info> seg_info;
    std::vector<int32_t> base;
    exc_lrc_info lrc{1.0f, 0, 0.0f};
    exc_seg_info list{1.0f, 0.0f, 0, 0};
    exc_gu_regression_info greg{};
    bool ok = true;
};

void encode_list(int codec, uint32_t seg, float frac, uint32_t hs_target, uint32_t hs_shift,
                 const uint32_t* L, uint32_t n, ListOut& o)
{
    if (codec == EXC_RAW || codec == EXC_RAW_HS) {  // the docIDs as they are (the GU programs' .ind2)
        o.words.assign(L, L + n);
        if (codec == EXC_RAW) o.greg = gu_regression(L, n);
        if (codec == EXC_RAW_HS) gu_hash_table(L, n, hs_target, hs_shift, o.hash);
        return;
    }
    std::vector<uint32_t> vd(n ? n : 1);
    if (codec == EXC_LRC) {
        o.lrc = lrc_fit(L, n);
        if (n > 1) {
            int32_t mn = 0;
            for (uint32_t i = 0; i < n; ++i) {
                int32_t d = (int32_t)(L[i] - lrc_pred(i, o.lrc.fSlope, o.lrc.iIntercept));
                vd[i] = (uint32_t)d;
                if (i == 0 || d < mn) mn = d;
            }
            for (uint32_t i = 0; i < n; ++i) vd[i] = (uint32_t)((int32_t)vd[i] - mn);
            o.lrc.iIntercept += mn;
        } else {
            vd[0] = 0;
        }
    }
    if ((codec == EXC_PARAPFD || codec == EXC_NEWPFD) && n) {
        vd[0] = L[0];
        for (uint32_t j = 1; j < n; ++j) vd[j] = L[j] - L[j - 1];
    }
    if (codec == EXC_SEGLRC) o.list = seg_fit(L, n);
    if (codec == EXC_HS_SEGLRC) {
        hs_table(L, n, hs_target, hs_shift, o.hash);
        uint32_t nb = o.hash[1];
        o.seg_offset.resize(nb);
        o.seg_info.resize(nb);
        o.base.resize(nb);
        for (uint32_t b = 0; b < nb; ++b) {
            uint32_t st = o.hash[2 + b], len = o.hash[3 + b] - st;
            o.seg_info[b] = seg_fit(L + st, len);
            o.base[b] = seg_residuals(L + st, len, o.seg_info[b], vd.data() + st);
            o.seg_offset[b] = (uint32_t)o.words.size();
            if (len && !pfor_block(vd.data() + st, len, frac, o.words)) { o.ok = false; return; }
        }
        return;
    }
    uint32_t nseg = (n + seg - 1) / seg;
    o.seg_offset.resize(nseg);
    o.seg_head.resize(nseg);
    if (codec == EXC_PARAPFD || codec == EXC_NEWPFD) o.seg_median.resize(nseg);
    if (codec == EXC_SEGLRC) { o.seg_info.resize(nseg); o.base.resize(nseg); }
    for (uint32_t s = 0; s < nseg; ++s) {
        uint32_t st = s * seg, len = std::min(seg, n - st);
        o.seg_head[s] = L[st];
        if (codec == EXC_PARAPFD || codec == EXC_NEWPFD) o.seg_median[s] = L[st + len / 2];
        if (codec == EXC_SEGLRC) {
            o.seg_info[s] = seg_fit(L + st, len);
            o.base[s] = seg_residuals(L + st, len, o.seg_info[s], vd.data() + st);
        }
        o.seg_offset[s] = (uint32_t)o.words.size();
        const bool ok = codec == EXC_NEWPFD ? npfd_block(vd.data() + st, len, frac, o.words)
                                            : pfor_block(vd.data() + st, len, frac, o.words);
        if (!ok) { o.ok = false; return; }
    }
}

template <class F>
void parallel_for(uint32_t n, int nthreads, F f)
{
    if (nthreads <= 1 || n < 2) { for (uint32_t i = 0; i < n; ++i) f(i, 0); return; }
    std::atomic<uint32_t> next{0};
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t)
        th.emplace_back([&, t] {
            for (;;) {
                uint32_t i = next.fetch_add(64);
                if (i >= n) break;
                for (uint32_t k = i; k < std::min(n, i + 64); ++k) f(k, t);
            }
        });
    for (auto& x : th) x.join();
}

}  // namespace

// A SegLRC index in the reference's arrays from device-resident records
// (exc_index_export of an index built on the GPU).
int exc::built_from_device(int codec, uint32_t seg, uint32_t nterms, const TermRec* terms, uint64_t nseg,
                           const SegRec* rec, const uint32_t* head, const uint32_t* words, uint64_t nwords,
                           const uint32_t* hash, uint64_t hash_words, exc_built** out)
{
    const bool hs = codec == EXC_HS_SEGLRC;
    auto* b = new exc_built();
    b->codec = codec;
    b->seg = seg;
    b->ind1.resize(nterms);
    b->seg_num.resize(nterms);
    if (codec == EXC_SEGLRC) b->list_info.resize(nterms);
    if (codec == EXC_LRC) b->lrc.resize(nterms);
    for (uint32_t t = 0; t < nterms; ++t) {
        const TermRec& R = terms[t];
        b->ind1[t].m_urlcount = R.n;
        b->ind1[t].m_off = R.woff * 4;
        b->seg_num[t] = R.segnum;
        if (codec == EXC_SEGLRC) {
            exc_seg_info& li = b->list_info[t];
            li.fSlope = R.slope;
            std::memcpy(&li.fIntercept, &R.icpt, 4);
            li.nRangeLeft = (uint32_t)R.range_l;
            li.nRangeRight = (uint32_t)R.range_r;
        } else if (codec == EXC_LRC) {
            b->lrc[t] = exc_lrc_info{R.slope, R.icpt, R.range_l};
        }
    }
    b->ind2.assign(words, words + nwords);
    b->seg_offset.resize(nseg);
    if (!hs) b->seg_head.assign(head, head + nseg);
    if (codec == EXC_SEGLRC || hs) {
        b->seg_info.resize(nseg);
        b->base.resize(nseg);
    }
    if (hs) {  // the tables, and each list's table offset (its record points past the header)
        b->hash_info.assign(hash, hash + hash_words);
        b->hash_offset.resize(nterms);
        for (uint32_t t = 0; t < nterms; ++t) b->hash_offset[t] = terms[t].aux - 2;
    }
    if (codec == EXC_PARAPFD) b->seg_median.resize(nseg);
    for (uint64_t g = 0; g < nseg; ++g) {
        b->seg_offset[g] = rec[g].off;
        if (codec == EXC_SEGLRC || hs) {
            b->seg_info[g] = exc_seg_info{rec[g].slope, rec[g].icpt, rec[g].rl, rec[g].rr};
            b->base[g] = rec[g].base;
        }
        if (codec == EXC_PARAPFD) b->seg_median[g] = rec[g].rl;  // the .segMedian row
    }
    *out = b;
    return EXC_OK;
}

extern "C" int exc_build_index(int codec, uint32_t seg, float frac, uint32_t hs_target, uint32_t hs_shift,
                               const uint32_t* docs, const uint64_t* list_off, uint32_t nterms,
                               int nthreads, exc_built** out)
{
    if (!out || !list_off || (nterms && !docs)) return set_error(EXC_ERR_ARG, "exc_build_index: null argument");
    if (codec < EXC_LRC || codec > EXC_RAW_HS) return set_error(EXC_ERR_ARG, "exc_build_index: unknown codec");
    if (codec == EXC_RAW || codec == EXC_RAW_HS) seg = 128;  // not stored: the unit planner's block size
    if (hs_shift > 31) return set_error(EXC_ERR_ARG, "exc_build_index: hash shift %u above 31", hs_shift);
    if (codec != EXC_HS_SEGLRC && (seg < 64 || seg > 1024 || (seg & 63)))
        return set_error(EXC_ERR_ARG, "exc_build_index: segment size must be a multiple of 64 in [64,1024]");
    if (hs_target == 0) hs_target = 256;
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    for (uint32_t t = 0; t < nterms; ++t) {
        uint64_t n = list_off[t + 1] - list_off[t];
        if (list_off[t + 1] < list_off[t] || n > 0xffffffffull)
            return set_error(EXC_ERR_ARG, "exc_build_index: bad list offsets");
        for (uint64_t i = 1; i < n; ++i)
            if (docs[list_off[t] + i] <= docs[list_off[t] + i - 1])
                return set_error(EXC_ERR_ARG, "exc_build_index: list %u is not strictly increasing", t);
        if (n && docs[list_off[t + 1] - 1] > 0x7fffffffu)
            return set_error(EXC_ERR_ARG, "exc_build_index: docIDs must be < 2^31");
    }
    std::vector<ListOut> lo(nterms);
    parallel_for(nterms, nthreads, [&](uint32_t t, int) {
        encode_list(codec, seg, frac, hs_target, hs_shift, docs + list_off[t],
                    (uint32_t)(list_off[t + 1] - list_off[t]), lo[t]);
    });
    for (uint32_t t = 0; t < nterms; ++t)
        if (!lo[t].ok) return set_error(EXC_ERR_FORMAT, "exc_build_index: list %u has an unrepresentable block", t);

    auto* b = new exc_built();
    b->codec = codec;
    b->seg = seg;
    b->ind1.resize(nterms);
    b->seg_num.resize(nterms);
    std::vector<uint64_t> woff(nterms + 1, 0), soff(nterms + 1, 0), hoff(nterms + 1, 0);
    for (uint32_t t = 0; t < nterms; ++t) {
        woff[t + 1] = woff[t] + lo[t].words.size();
        soff[t + 1] = soff[t] + lo[t].seg_offset.size();
        hoff[t + 1] = hoff[t] + lo[t].hash.size();
    }
    if (soff[nterms] > 0xffffffffull || hoff[nterms] > 0xffffffffull) {
        delete b;
        return set_error(EXC_ERR_CAPACITY, "exc_build_index: more than 2^32 segments");
    }
    uint64_t S = soff[nterms];
    b->ind2.resize(woff[nterms]);
    b->seg_offset.resize(S);
    if (codec != EXC_HS_SEGLRC) b->seg_head.resize(S);
    if (codec == EXC_PARAPFD || codec == EXC_NEWPFD) b->seg_median.resize(S);
    if (codec == EXC_SEGLRC || codec == EXC_HS_SEGLRC) { b->seg_info.resize(S); b->base.resize(S); }
    if (codec == EXC_SEGLRC) b->list_info.resize(nterms);
    if (codec == EXC_LRC) b->lrc.resize(nterms);
    if (codec == EXC_RAW) b->gu_reg.resize(nterms);
    const bool hashed = codec == EXC_HS_SEGLRC || codec == EXC_RAW_HS;
    if (hashed) { b->hash_offset.resize(nterms); b->hash_info.resize(hoff[nterms]); }
    parallel_for(nterms, nthreads, [&](uint32_t t, int) {
        ListOut& o = lo[t];
        b->ind1[t].m_urlcount = (uint32_t)(list_off[t + 1] - list_off[t]);
        b->ind1[t].m_off = woff[t] * 4;
        b->seg_num[t] = (uint32_t)soff[t];
        std::copy(o.words.begin(), o.words.end(), b->ind2.begin() + woff[t]);
        std::copy(o.seg_offset.begin(), o.seg_offset.end(), b->seg_offset.begin() + soff[t]);
        if (!o.seg_head.empty()) std::copy(o.seg_head.begin(), o.seg_head.end(), b->seg_head.begin() + soff[t]);
        if (!o.seg_median.empty()) std::copy(o.seg_median.begin(), o.seg_median.end(), b->seg_median.begin() + soff[t]);
        if (!o.seg_info.empty()) std::copy(o.seg_info.begin(), o.seg_info.end(), b->seg_info.begin() + soff[t]);
        if (!o.base.empty()) std::copy(o.base.begin(), o.base.end(), b->base.begin() + soff[t]);
        if (codec == EXC_SEGLRC) b->list_info[t] = o.list;
        if (codec == EXC_LRC) b->lrc[t] = o.lrc;
        if (codec == EXC_RAW) b->gu_reg[t] = o.greg;
        if (hashed) {
            b->hash_offset[t] = (uint32_t)hoff[t];
            std::copy(o.hash.begin(), o.hash.end(), b->hash_info.begin() + hoff[t]);
        }
        ListOut().words.swap(o.words);
    });
    *out = b;
    return EXC_OK;
}

template <class T>
static const T* ptr_or_null(const std::vector<T>& v) { return v.empty() ? nullptr : v.data(); }

extern "C" int exc_built_view(const exc_built* b, exc_index_host* v)
{
    if (!b || !v) return set_error(EXC_ERR_ARG, "exc_built_view: null argument");
    std::memset(v, 0, sizeof(*v));
    v->codec = b->codec;
    v->seg_size = b->seg;
    v->nterms = (uint32_t)b->ind1.size();
    v->ind1 = ptr_or_null(b->ind1);
    v->ind2 = ptr_or_null(b->ind2);
    v->ind2_words = b->ind2.size();
    v->seg_num = ptr_or_null(b->seg_num);
    v->seg_offset = ptr_or_null(b->seg_offset);
    v->nseg = b->seg_offset.size();
    v->seg_head = ptr_or_null(b->seg_head);
    v->seg_median = ptr_or_null(b->seg_median);
    v->lrc_info = ptr_or_null(b->lrc);
    v->seg_info = ptr_or_null(b->seg_info);
    v->seg_base = ptr_or_null(b->base);
    v->list_info = ptr_or_null(b->list_info);
    v->hash_offset = ptr_or_null(b->hash_offset);
    v->hash_info = ptr_or_null(b->hash_info);
    v->hash_words = b->hash_info.size();
    v->gu_regression = ptr_or_null(b->gu_reg);
    return EXC_OK;
}

extern "C" void exc_built_free(exc_built* b) { delete b; }

// ------------------------------------------------------------ index files

static int write_file(const std::string& p, const void* data, size_t bytes)
{
    FILE* f = std::fopen(p.c_str(), "wb");
    if (!f) return set_error(EXC_ERR_IO, "cannot create %s", p.c_str());
    size_t w = bytes ? std::fwrite(data, 1, bytes, f) : 0;
    std::fclose(f);
    return w == bytes ? EXC_OK : set_error(EXC_ERR_IO, "short write on %s", p.c_str());
}

extern "C" int exc_index_save_files(const exc_index_host* h, const char* dir, const char* name)
{
    if (!h || !dir || !name) return set_error(EXC_ERR_ARG, "exc_index_save_files: null argument");
    std::string p = std::string(dir) + "/" + name;
    int rc;
    uint32_t nt = h->nterms;
    if ((rc = write_file(p + ".ind1", h->ind1, (size_t)nt * sizeof(exc_term_ind1)))) return rc;
    if ((rc = write_file(p + ".ind2", h->ind2, h->ind2_words * 4))) return rc;
    if ((rc = write_file(p + ".segNum", h->seg_num, (size_t)nt * 4))) return rc;
    if ((rc = write_file(p + ".segOffset", h->seg_offset, h->nseg * 4))) return rc;
    if (h->seg_head && (rc = write_file(p + ".segHead", h->seg_head, h->nseg * 4))) return rc;
    if (h->seg_median && (rc = write_file(p + ".segMedian", h->seg_median, h->nseg * 4))) return rc;
    if (h->lrc_info && (rc = write_file(p + ".regression_info", h->lrc_info, (size_t)nt * sizeof(exc_lrc_info)))) return rc;
    if (h->seg_info && (rc = write_file(p + ".regression_info", h->seg_info, h->nseg * sizeof(exc_seg_info)))) return rc;
    if (h->list_info && (rc = write_file(p + ".regression_info_list", h->list_info, (size_t)nt * sizeof(exc_seg_info)))) return rc;
    if (h->seg_base && (rc = write_file(p + ".ind2_distance_base", h->seg_base, h->nseg * 4))) return rc;
    // the compressed HS programs name the tables <name>.HashSeg_*, the
    // uncompressed one <name>_HashSeg_* (cuda_hashSegment.cu:158,175, "at_")
    const std::string hp = p + (h->codec == EXC_RAW_HS ? "_HashSeg" : ".HashSeg");
    if (h->hash_offset && (rc = write_file(hp + "_offset", h->hash_offset, (size_t)nt * 4))) return rc;
    if (h->hash_info && (rc = write_file(hp + "_info", h->hash_info, h->hash_words * 4))) return rc;
    // the regression program's at_regression_info (cuda_linearRegression.cu:888)
    if (h->gu_regression &&
        (rc = write_file(p + "_regression_info", h->gu_regression, (size_t)nt * sizeof(exc_gu_regression_info))))
        return rc;
    return EXC_OK;
}

// ------------------------------------------------------------ synthetic lists

namespace {

struct SplitMix {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // uniform in [0, bound) by 64x64 multiply-high (Lemire)
    uint32_t below(uint32_t bound) { return (uint32_t)(((unsigned __int128)next() * bound) >> 64); }
};

void synth_one(uint64_t seed, uint32_t U, uint32_t t, uint32_t df, uint32_t* out,
               std::vector<uint64_t>& bitmap, std::vector<uint32_t>& tmp)
{
    SplitMix r{seed * 0x100000001B3ull ^ (0x9E3779B97F4A7C15ull * (uint64_t)(t + 1))};
    if (df == 0) return;
    if ((uint64_t)df * 256 >= U) {
        // dense: Floyd's sampling over a bitmap, read back in order
        bitmap.assign((U + 63) / 64, 0);
        for (uint32_t j = U - df; j < U; ++j) {
            uint32_t x = r.below(j + 1);
            if (bitmap[x >> 6] >> (x & 63) & 1) x = j;
            bitmap[x >> 6] |= 1ull << (x & 63);
        }
        uint32_t k = 0;
        for (uint32_t w = 0; w < bitmap.size(); ++w)
            for (uint64_t m = bitmap[w]; m; m &= m - 1) out[k++] = w * 64 + (uint32_t)__builtin_ctzll(m);
        return;
    }
    // sparse: draw, sort, dedupe, top up until df distinct values
    tmp.clear();
    while (tmp.size() < df) {
        size_t need = df - tmp.size();
        for (size_t i = 0; i < need; ++i) tmp.push_back(r.below(U));
        std::sort(tmp.begin(), tmp.end());
        tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    }
    std::copy(tmp.begin(), tmp.end(), out);
}

}  // namespace

extern "C" int exc_synth_lists(uint64_t seed, uint32_t universe, const uint32_t* df, uint32_t nterms,
                               int nthreads, uint32_t* docs, const uint64_t* list_off)
{
    if (!df || !docs || !list_off || universe == 0) return set_error(EXC_ERR_ARG, "exc_synth_lists: bad argument");
    for (uint32_t t = 0; t < nterms; ++t)
        if (df[t] > universe || list_off[t + 1] - list_off[t] != df[t])
            return set_error(EXC_ERR_ARG, "exc_synth_lists: df/list_off mismatch at term %u", t);
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    std::vector<std::vector<uint64_t>> bm(nthreads);
    std::vector<std::vector<uint32_t>> tmp(nthreads);
    parallel_for(nterms, nthreads, [&](uint32_t t, int w) {
        synth_one(seed, universe, t, df[t], docs + list_off[t], bm[w], tmp[w]);
    });
    return EXC_OK;
}
