// exc_internal.h -- device-side index layout and kernel launch interface shared
// by the runtime (runtime.cpp) and the gfx950 kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/exclucene_amd.h"

int set_error(int code, const char* fmt, ...);

namespace exc {

// One record per term, built at upload from .ind1/.segNum/.regression_info(_list)
// /.HashSeg_offset, so a wave fetches everything it needs about a list with one
// 64-byte scalar load.
struct alignas(16) TermRec {
    uint64_t woff;    // word offset of the list's first block in ind2
    uint32_t n;       // list length (ind1.m_urlcount)
    uint32_t segnum;  // first segment record (.segNum; HS: first record of a non-empty bucket)
    uint32_t nseg;    // segments (HS: non-empty buckets, one record each; RAW_HS: buckets)
    uint32_t aux;     // HS: word offset of the list's bucket-occupancy table in
                      // DevIndex::hash, or kDenseBuckets (every bucket holds docIDs);
                      // RAW_HS: index of start_0 in .HashSeg_info
    float slope;      // LRC fSlope | SegLRC list-level / RAW regression slope
    int32_t icpt;     // LRC iIntercept | SegLRC list-level / RAW regression intercept bits | HS shift
    float range_l;    // LRC fSearchRange | SegLRC list-level / RAW regression nRangeLeft
    float range_r;    // SegLRC list-level / RAW regression nRangeRight
    uint32_t model;   // learned model slot, or kNoModel
    uint32_t term;    // the term's own index
    float merr;       // learned model: recorded max error (positions)
    float mx0;        //   input offset
    float mxs;        //   input scale
    uint32_t mkind;   //   kModelPla | kModelMlp
};
static_assert(sizeof(TermRec) == 64, "TermRec is one 64-byte record");

constexpr uint32_t kNoModel = 0xffffffffu;

// HS_SegLRC on the device keeps a record only for buckets that hold docIDs: a
// fixed shift leaves most buckets of a short list empty, and a 32-byte record
// per empty bucket is most of such an index.  A list with empty buckets has an
// occupancy table [nb, 0, then per 32 buckets {bits, records before them}]
// (8-byte aligned pairs), so bucket b's record is
// segnum + prefix[b / 32] + popcount(bits[b / 32] & ((1 << b % 32) - 1)),
// present iff bit b % 32 is set; a list whose buckets all hold docIDs has no
// table (aux = kDenseBuckets) and record b is bucket b.  DevIndex::seg_head
// holds each record's first element index, densely, for the chunk decoder.
constexpr uint32_t kDenseBuckets = 0xffffffffu;
constexpr uint32_t hs_occ_words(uint32_t nb) { return 2u + 2u * ((nb + 31u) / 32u); }

// Learned position model of one list (docID -> position), fp32; inference in
// exc_model.h.  err = max |prediction - true position| over the list's own
// elements, measured on the device with the same inference code the search uses.
struct alignas(32) Model {
    uint32_t kind;       // kModelPla | kModelMlp
    uint32_t n;          // list length
    float err;           // recorded max error (positions)
    float x0;            // input offset (first docID)
    float xs;            // input scale
    uint32_t param_off;  // float offset into the parameter pool
    uint32_t term;       // the term the model belongs to
    uint32_t pad;
};
static_assert(sizeof(Model) == 32, "Model is 32 bytes");

// Per-segment (per-bucket) record derived at upload so one 32-byte load gives
// a probe everything about a block: its word offset, ParaPFor descriptor,
// fit, base and first docID (the .segOffset/.regression_info/
// .ind2_distance_base/.segHead rows of the reference, interleaved).
struct alignas(32) SegRec {
    uint32_t off;    // word offset of the block inside its list
    uint32_t desc;   // the block's ParaPFor descriptor word
    float slope;     // SegLRC/HS fit slope (1 otherwise)
    float icpt;      // SegLRC/HS fit intercept
    uint32_t rl;     // nRangeLeft
    uint32_t rr;     // nRangeRight
    int32_t base;    // SegLRC/HS residual base
    uint32_t head;   // first docID (segment codecs) | first element index (HS)
};
static_assert(sizeof(SegRec) == 32, "SegRec is 32 bytes");

struct DevIndex {
    const SegRec* rec;
    const uint32_t* words;
    const TermRec* terms;
    const uint32_t* seg_head;  // .segHead (dense, for the head searches); HS: first element of each record
    const uint32_t* hash;      // RAW_HS: .HashSeg_info; HS: the bucket-occupancy tables
    const Model* models;
    const float* mparams;
    int32_t codec;
    uint32_t seg;
    uint32_t flags;  // kIxMedians | kIxQuarters | kIxOctants
    const uint32_t* pfd_oct;  // ParaPFD (kIxOctants): per block the values at pfd_oct_pos(len, 0..3)
};
// DevIndex::flags: ParaPFD records carry the block medians (.segMedian) in rl
constexpr uint32_t kIxMedians = 1;
// ... and the values at the block's quarter positions pfd_m1 (in base) and
// pfd_m3 (in rr), derived on the device when the index is built or uploaded
constexpr uint32_t kIxQuarters = 2;
// ... and the values at the midpoints of the quarters in pfd_oct
constexpr uint32_t kIxOctants = 4;

// One ParaPFD block whose quarter values k_pfd_anchors takes from a decoded
// run: its record, its first value's position in the run and its length.
struct AnchorJob {
    uint64_t rec;
    uint64_t pos;
    uint32_t len;
    uint32_t pad;
};

struct QueryRec {
    uint32_t toff;        // first sorted term of the query in qterms
    uint32_t tc;          // term count
    uint32_t first_unit;  // first work unit of the query
    uint32_t nunits;      // work units (shortest-list chunks)
    uint64_t short_off;   // offset of the query's shortest list in the stage buffer
    uint32_t knobs;       // search knobs of this query, see make_knobs
    uint32_t klass;       // list-length-ratio class (kRatioClasses)
};
static_assert(sizeof(QueryRec) == 32, "QueryRec is 32 bytes");

enum { SEARCH_BINARY = 0, SEARCH_FIT = 1, SEARCH_LEARNED = 2, SEARCH_INTERP = 3 };

// Per-query search knobs: bits 0-7 = longest run of a longer list's segments a
// wave decodes cooperatively (0 = one lane per candidate, 64 = always
// cooperative), bits 8-9 = search mode.
constexpr uint32_t make_knobs(uint32_t coop, uint32_t search) { return (coop & 0xffu) | ((search & 3u) << 8); }
__host__ __device__ __forceinline__ uint32_t knob_coop(uint32_t k) { return k & 0xffu; }
__host__ __device__ __forceinline__ uint32_t knob_search(uint32_t k) { return (k >> 8) & 3u; }

// List-length-ratio classes: class c < 4 holds queries whose second-shortest
// list is 10^c .. 10^(c+1) times the shortest, class 4 the ratios >= 10^4,
// class 5 single-term queries (nothing to probe).
constexpr int kRatioClasses = 6;
static_assert(kRatioClasses == EXC_RATIO_CLASSES, "ratio classes");

// One record per work unit (query, kChunk-docID slice k of its shortest
// list), carrying what a wave needs before its first index read.
struct UnitRec {
    uint32_t q;      // query
    uint32_t k;      // slice of the shortest list
    uint32_t toff;   // the query's first term record in the batch copy (sorted by length)
    uint16_t tc;     // term count
    uint16_t knobs;  // search knobs (make_knobs)
    // Filled on the device before the probe (k_unit_prologue): the blocks of
    // the second-shortest list that can hold a docID of this slice, found from
    // the slice's bounding segment heads.  l1_count = kUnset: not computed,
    // 0: no docID of the slice can be in that list.
    uint32_t l1_first;
    uint32_t l1_count;
    // index of this unit in query order (its count and stage-scan slot): the
    // probe runs units in launch order, the scan in query order
    uint32_t slot;
    // stage-buffer offset of this unit's survivors (the query's shortest-list
    // offset + k * kChunk; the planner keeps the summed length below 2^32)
    uint32_t soff;
};
static_assert(sizeof(UnitRec) == 32, "UnitRec is 32 bytes");
constexpr uint32_t kUnset = 0xffffffffu;

struct ProbeArgs {
    DevIndex ix;
    const QueryRec* queries;
    const TermRec* qrecs;  // per query term, sorted by length: copies of the index's TermRecs
    UnitRec* units;
    uint32_t nunits;
    int32_t persistent;
    uint32_t* stage;
    uint32_t* ucount;
    uint32_t* counter;
};

struct ScanArgs {
    const QueryRec* queries;
    const UnitRec* units;
    const uint32_t* ucount;
    const uint32_t* stage;
    uint32_t nunits;
    uint32_t nq;
    uint32_t pad_;
    uint32_t* uoff;     // per unit exclusive offset inside its scan block
    uint32_t* bsum;     // per scan block sum, then exclusive block offsets (+ total)
    uint32_t* result;
    uint32_t* qcum;     // per query inclusive cumulative count
    uint32_t* ticket;   // blocks of k_scan_local done (zero between launches)
};

constexpr uint32_t kScanBlockItems = 4096;  // units per scan block (256 threads x 16)
constexpr uint32_t kScatterUnits = 8;       // consecutive units per wave of k_scatter
constexpr uint32_t kScatterRows = 8;        // 64-element rows k_scatter reads before writing
constexpr uint32_t kScanTicketWord = 32;     // word of the context's counter buffer holding the scan ticket
constexpr uint32_t kChunk = 512;            // shortest-list docIDs per work unit
constexpr uint32_t kDecodeUnit = 1024;      // docIDs per wave of whole-list decompression (HS: one chunk)

// Batch planning on the device, for the entry point that allocates nothing
// (exc_intersect_noalloc): the query list is copied into a fixed device
// scratch area and these kernels write the records plan_queries writes on the
// host -- terms sorted by ascending length (stable), their TermRec copies,
// one QueryRec per query, the work units in query order.
constexpr uint32_t kNaMaxTerms = 64;  // terms per query the device planner sorts
struct NaPlanArgs {
    const TermRec* terms;    // the index's term records
    const uint32_t* qterms;  // the batch's terms as given
    const uint64_t* qoff;    // nq + 1 offsets into qterms (qoff[0] need not be 0)
    uint32_t nq;
    uint32_t knobs[kRatioClasses];  // search knobs per ratio class
    QueryRec* queries;
    TermRec* qrecs;          // one per batch term, sorted per query
    UnitRec* units;          // query order
};
hipError_t launch_noalloc_plan(const NaPlanArgs& a, hipStream_t s);
hipError_t launch_zero_words(uint32_t* p, uint32_t n, hipStream_t s);

// launchers (kernels.hip); all return hipError_t of the launch
hipError_t launch_reset(uint32_t* counter, hipStream_t s);
hipError_t launch_probe(const ProbeArgs& a, uint32_t grid, hipStream_t s);
hipError_t launch_unit_prologue(const ProbeArgs& a, hipStream_t s);

// ---- device index builder (gpubuild.hip)
// Synthetic lists: item i = chunk item_chunk[i] of term item_term[i]'s docID
// range (nchunks[t] equal spans of [0, universe)); docIDs of item i start at
// item_off[i] of the virtual concatenation of all lists, docs[] holds the
// part from doc_base on.
struct SynthArgs {
    uint64_t seed;
    uint32_t universe;
    uint32_t pad_;
    const uint32_t* df;
    const uint32_t* nchunks;
    const uint32_t* item_term;
    const uint32_t* item_chunk;
    const uint64_t* item_off;
    uint32_t* counts;
    uint32_t* docs;
    uint64_t doc_base;
    uint64_t item0, item1;
};
// Encoding of segments [g0, g1) and lists [t0, t1) whose docIDs are in docs[]
// from doc_base on (list t at list_off[t] of the concatenation): SegLRC, LRC
// (residuals against the list fit) or ParaPFD (d-gaps), 128- or 256-docID blocks.
struct EncodeArgs {
    const uint32_t* docs;
    const uint64_t* list_off;  // nterms + 1
    uint64_t doc_base;
    const uint32_t* seg_num;   // nterms + 1
    uint32_t nterms;
    uint32_t seg;
    uint64_t g0, g1;
    uint32_t t0, t1;
    SegRec* rec;
    uint32_t* head;
    uint32_t* nwords;          // per segment: descriptor + packed words
    const uint64_t* goff;      // per segment (+1): global word offset
    uint32_t* words;
    TermRec* terms;
    int32_t codec;             // EXC_SEGLRC | EXC_LRC | EXC_PARAPFD | EXC_HS_SEGLRC
    float frac;                // ParaPFor exception fraction (ParaPFD)
    uint32_t* err;             // set non-zero when a block cannot be encoded
    // HS_SegLRC: per list its shift and bucket count (first pass), the word
    // offset of its [shift][nb][start_0 .. start_nb] table, and the tables
    uint32_t* hs_shift;
    uint32_t* hs_nb;
    const uint64_t* hs_off;
    uint32_t* hash;
    uint32_t hs_target, hs_fixed;
    uint32_t* pfd_oct;         // ParaPFD: 4 values a block at pfd_oct_pos (nullptr: none)
};
hipError_t launch_synth(const SynthArgs& a, bool write, hipStream_t s);
hipError_t launch_seg_fit(const EncodeArgs& a, hipStream_t s);
hipError_t launch_seg_pack(const EncodeArgs& a, hipStream_t s);
// LRC / ParaPFD: list fits (LRC), then per block its plan (descriptor, words,
// head, median), then after the offset scan its words
hipError_t launch_lrc_lists(const EncodeArgs& a, hipStream_t s);
hipError_t launch_blk_fit(const EncodeArgs& a, hipStream_t s);
hipError_t launch_blk_pack(const EncodeArgs& a, hipStream_t s);
// HS_SegLRC: per list its shift and bucket count, then per bucket its table
// entry, fit, base and size, then after the offset scan its words
hipError_t launch_hs_lists(const EncodeArgs& a, hipStream_t s);
hipError_t launch_hs_fit(const EncodeArgs& a, hipStream_t s);
hipError_t launch_hs_pack(const EncodeArgs& a, hipStream_t s);
hipError_t launch_term_recs(const EncodeArgs& a, hipStream_t s);
hipError_t launch_heads_line(const EncodeArgs& a, hipStream_t s);
hipError_t launch_pfd_anchors(const AnchorJob* jobs, uint32_t n, const uint32_t* dec, SegRec* rec, uint32_t* oct,
                              hipStream_t s);
// HS_SegLRC: the per-bucket records and tables of the encoder compacted to the
// device layout (records of non-empty buckets, occupancy tables; kDenseBuckets)
struct HsCompactArgs {
    const uint32_t* seg_num;  // nterms + 1: first bucket of each list, all buckets counted
    const uint64_t* hs_off;   // per list: word offset of its [shift][nb][start_0 .. start_nb] table
    const uint32_t* hash;     // those tables
    const SegRec* rec;        // per bucket
    uint32_t nterms;
    uint64_t nbuckets;
    uint32_t* flag;           // per bucket (+1): 1 if it holds docIDs
    const uint64_t* cidx;     // per bucket (+1): records before it (exclusive scan of flag)
    uint32_t* count;          // per list: records
    const uint64_t* occ_off;  // per list: word offset of its occupancy table, ~0 for a dense list
    const uint64_t* grp_off;  // per list (+1): occupancy groups of 32 buckets before it
    uint64_t ngroups;
    SegRec* crec;
    uint32_t* chead;
    uint32_t* occ;
    TermRec* terms;
};
hipError_t launch_hs_compact_flags(const HsCompactArgs& a, hipStream_t s);  // flag, then count after the scan
hipError_t launch_hs_compact_counts(const HsCompactArgs& a, hipStream_t s);
hipError_t launch_hs_compact(const HsCompactArgs& a, hipStream_t s);        // records, heads, tables, TermRecs
hipError_t scan_u32_to_u64(const uint32_t* in, uint64_t* out, uint64_t n, hipStream_t s);
int built_from_device(int codec, uint32_t seg, uint32_t nterms, const TermRec* terms, uint64_t nseg, const SegRec* rec,
                      const uint32_t* head, const uint32_t* words, uint64_t nwords, const uint32_t* hash,
                      uint64_t hash_words, exc_built** out);
hipError_t launch_scan(const ScanArgs& a, hipStream_t s);
// One whole-list decode work unit (kDecodeUnit docIDs of one list), with the
// list fields the decoder needs so that its first load brings all of them.
struct alignas(16) DecUnit {
    uint64_t out;     // output offset of the unit's first docID
    uint64_t woff;    // the list's word offset
    uint32_t segnum;  // the list's first segment
    uint32_t n;       // list length
    uint32_t k;       // first block of the unit (HS: chunk)
    uint32_t term;    // term index (the HS decoder reads its bucket table)
    float slope;      // LRC list-level fit
    int32_t icpt;
    uint32_t wend;    // the list's words in .ind2 (kUnset: unknown)
    uint32_t pad;
};
static_assert(sizeof(DecUnit) == 48, "DecUnit is 48 bytes");
hipError_t launch_decompress(const DevIndex& ix, const DecUnit* units, uint32_t n, uint32_t* out, hipStream_t s);
uint32_t probe_waves_per_block();

}  // namespace exc

namespace exc {
// learned models (learned.hip)
struct TrainArgs {
    const uint32_t* dec;       // decoded lists, concatenated
    const uint64_t* dec_off;   // per model: offset of its list in dec
    Model* models;             // per model (header written by the fit kernels)
    float* params;             // parameter pool
    uint32_t* err_bits;        // per model: max error as float bits (atomicMax)
    uint32_t nmodels;
    uint32_t epochs;           // MLP: Adam epochs (0: gradient only)
    uint32_t samples;          // MLP: samples per list
    float lr;
    uint64_t seed;
    float* grad_out;           // MLP gradient-only mode: per model kMlpParams floats + loss
};
hipError_t launch_pla_fit(const TrainArgs& a, hipStream_t s);
hipError_t launch_plaq_fit(const TrainArgs& a, hipStream_t s);
hipError_t launch_mlp_train(const TrainArgs& a, hipStream_t s);
hipError_t launch_model_err(const TrainArgs& a, uint64_t total_elems, hipStream_t s);

// persistent-launch sizing: resident probe workgroups per CU for this codec / segment size
int probe_blocks_per_cu(int codec, uint32_t seg);
}  // namespace exc
