/*
 * exclucene_amd.h -- C ABI of the MI355X (gfx950) posting-list decode +
 * intersection library.
 *
 * The reference (ExCLucene src/gpu) has no library: each codec is a standalone
 * program whose host functions share file-scope globals.  This header replaces
 * those per-program functions with one context-handle API over the same
 * compressed index layouts.  Every entry point names the reference interface it
 * replaces (paths relative to the reference root; GC = src/gpu/Compressed/src).
 *
 * Conventions: plain pointers and sizes, no exit() -- every call returns an
 * EXC_* status and exc_last_error() holds a message for the calling thread.
 * One exc_ctx per host thread (a context owns one HIP stream and its buffers).
 */
#ifndef EXCLUCENE_AMD_H
#define EXCLUCENE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference calls exit(); LRC_Intersection_kernel.cu:7-14) */
#define EXC_OK            0
#define EXC_ERR_ARG      (-1)  /* bad argument / malformed query stream */
#define EXC_ERR_DEVICE   (-2)  /* HIP runtime error */
#define EXC_ERR_CAPACITY (-3)  /* batch or output larger than the buffers (LRC_Intersection.cu:218-226) */
#define EXC_ERR_FORMAT   (-4)  /* index arrays inconsistent with the codec's layout */
#define EXC_ERR_IO       (-5)  /* index file missing / short */
#define EXC_ERR_BUDGET   (-6)  /* device-memory budget exceeded */
#define EXC_ERR_STATE    (-7)  /* call out of order (e.g. dtoh before kernelInvoke) */
#define EXC_ERR_KERNEL   (-8)  /* a kernel reported an internal limit (bounded wait expired) */

/* ---- codecs: one per reference program family */
enum {
    EXC_LRC = 0,       /* GC/LRC       : one global fit per list          */
    EXC_SEGLRC = 1,    /* GC/SegLRC    : per-segment fit + base           */
    EXC_HS_SEGLRC = 2, /* GC/HS_SegLRC : hash buckets on top docID bits   */
    EXC_PARAPFD = 3,   /* GC/ParaPFD   : d-gaps in ParaPFor blocks        */
    EXC_NEWPFD = 4,    /* GC/NewPFD    : d-gaps in NewPFor blocks, Simple16
                          exceptions; decompression only, as in the reference
                          (GC/NewPFD/NewPFD_Decompression)                  */
    EXC_RAW = 5,       /* GU (src/gpu/Uncompressed): .ind2 holds the sorted
                          docIDs themselves; EXC_OPT_SEARCH 0 = binary search
                          (GU/cuda_binarySearch), 1 = the list's regression
                          window (GU/cuda_runRegressionAlgorithm), 2 = its
                          learned window, 3 = interpolation search
                          (GU/cuda_interpolationSearch)                      */
    EXC_RAW_HS = 6     /* GU/cuda_hashSegment_new: the sorted docIDs as in
                          EXC_RAW plus per-list bucket tables on the top docID
                          bits, [shift][start_0 .. start_nb] in
                          <name>_HashSeg_info, word offsets in
                          <name>_HashSeg_offset (cuda_hashSegment.cu:815-947);
                          a probe searches only its docID's bucket
                          (cuda_hashSegment_kernel.cu:269-307), bound-checked */
};

/* ---- reference structs, byte-identical */
/* .ind1 row (GC/LRC/LRC_Compression/main.cpp:34-37), 16 bytes with padding. */
typedef struct { uint32_t m_urlcount; uint64_t m_off; } exc_term_ind1;
/* LRC .regression_info row (GC/LRC/LRC_Intersection/LRC_Intersection_kernel.cu:89-93). */
typedef struct { float fSlope; int32_t iIntercept; float fSearchRange; } exc_lrc_info;
/* SegLRC/HS .regression_info row (GC/SegLRC/SegLRC_Intersection/SegLRC_Intersection_kernel.cu:89-94). */
typedef struct { float fSlope; float fIntercept; uint32_t nRangeLeft; uint32_t nRangeRight; } exc_seg_info;
/* Uncompressed-list regression row, the at_regression_info record of
 * GU/cuda_runRegressionAlgorithm (cuda_linearRegression_kernel.cu:105-120),
 * 36 bytes: the list's least-squares line over (position, docID) and the
 * furthest positions on either side of it (cuda_linearRegression.cu:792-873).
 * The probe reads fSlope, fIntercept, nRangeLeft and nRangeRight. */
typedef struct {
    float fSlope, fIntercept, fRSquare, fMultiple, fInterceptHigh, fInterceptLow;
    uint32_t nRangeLeft, nRangeRight, nMagnification;
} exc_gu_regression_info;
/* batchInfo (GC/LRC/LRC_Intersection/LRC_Intersection_kernel.cu:68-74). */
typedef struct {
    uint32_t blockNum;           /* work units (query, shortest-list segment) in the batch */
    uint32_t constantUsedInByte; /* bytes of batch descriptor sent to the device           */
    uint32_t ucntShortest_sum;   /* sum of shortest-list lengths                           */
    uint32_t nTotalQueryNum;     /* queries in the batch                                   */
    uint32_t nLinearInfoInByte;  /* bytes of per-term records referenced by the batch      */
} exc_batchInfo;
/* checkSum (LRC_Intersection_kernel.cu:79-83). */
typedef struct { uint64_t checkSum1, checkSum2, checkSum3; } exc_checkSum;

/* Host view of one compressed index, array for array the reference's files:
 * .ind1 .ind2 .segNum .segOffset .segHead .segMedian .regression_info
 * .regression_info_list .ind2_distance_base .HashSeg_offset .HashSeg_info
 * (GC/<codec>/<codec>_Compression/main.cpp, GC/<codec>/<codec>_Generator).  Arrays a codec does not
 * use are NULL.  seg_size is the segment length (the reference fixes 256). */
typedef struct {
    int32_t codec;
    uint32_t seg_size;
    uint32_t nterms;
    uint32_t reserved;
    const exc_term_ind1 *ind1;     /* nterms; m_off in bytes into ind2            */
    const uint32_t *ind2;          /* compressed words                            */
    uint64_t ind2_words;
    const uint32_t *seg_num;       /* nterms: first segment / bucket of the term  */
    const uint32_t *seg_offset;    /* nseg: block word offset inside its list     */
    uint64_t nseg;
    const uint32_t *seg_head;      /* nseg: first docID (LRC, SegLRC, ParaPFD)    */
    const uint32_t *seg_median;    /* nseg: ParaPFD only (optional)               */
    const exc_lrc_info *lrc_info;  /* nterms: LRC                                 */
    const exc_seg_info *seg_info;  /* nseg: SegLRC / HS                           */
    const int32_t *seg_base;       /* nseg: SegLRC / HS (.ind2_distance_base)     */
    const exc_seg_info *list_info; /* nterms: SegLRC list-level fit (optional)    */
    const uint32_t *hash_offset;   /* nterms: HS, RAW_HS                          */
    const uint32_t *hash_info;     /* HS tables [shift][nb][start_0..start_nb];
                                      RAW_HS tables [shift][start_0..start_nb]    */
    uint64_t hash_words;
    const exc_gu_regression_info *gu_regression; /* nterms: RAW (optional;
                                      <name>_regression_info on disk)             */
} exc_index_host;

typedef struct exc_ctx exc_ctx;
typedef struct exc_index exc_index;
typedef struct exc_built exc_built;

/* ---- context (replaces allocateResource/releaseResource globals,
 *      GC/LRC/LRC_Intersection/LRC_Intersection.cu:77-122,382-446) */
int exc_ctx_create(int device, uint64_t mem_budget_bytes, exc_ctx **out);
void exc_ctx_destroy(exc_ctx *ctx);
const char *exc_last_error(void);

/* Device buffers for callers without a HIP runtime of their own (no
 * reference counterpart; the reference allocates everything itself in
 * allocateResource).  They count against the context's budget and are freed
 * with the context at the latest.  exc_memcpy waits for the context's stream;
 * to_device = 1 copies host -> device, 0 device -> host. */
int exc_device_alloc(exc_ctx *ctx, uint64_t bytes, void **d_ptr);
int exc_device_free(exc_ctx *ctx, void *d_ptr);
int exc_memcpy(exc_ctx *ctx, void *dst, const void *src, uint64_t bytes, int to_device);
/* Page-locked host buffers (the reference's cudaMallocHost buffers of
 * allocateResource, GC/LRC/LRC_Intersection/LRC_Intersection.cu:81-82): a
 * result buffer here takes exc_dtohTransfer / exc_intersect's copy at full
 * PCIe rate instead of through the runtime's pageable staging.  Freed with
 * the context at the latest. */
int exc_host_alloc(exc_ctx *ctx, uint64_t bytes, void **h_ptr);
int exc_host_free(exc_ctx *ctx, void *h_ptr);

/* options (no reference counterpart: the reference fixes these with #define) */
enum {
    EXC_OPT_SEARCH = 0,     /* 0 plain binary search, 1 fit window + fallback, 2 learned window + fallback,
                               3 interpolation search (uncompressed lists; a compressed codec takes it as 1) */
    EXC_OPT_TPE = 1,        /* threads per short-list element: 1 random access, 64 wave-cooperative decode,
                               0 cooperative for runs up to EXC_OPT_COOP_MAX blocks.  Setting SEARCH, TPE
                               or COOP_MAX drops a table installed by exc_autotune/exc_ctx_set_tuning. */
    EXC_OPT_PERSISTENT = 2, /* 1: persistent work-queue launch, 0: one wave per work unit */
    EXC_OPT_GRAPH = 3,      /* 1: replay kernelInvoke as a captured hipGraph */
    EXC_OPT_TIMING = 4,     /* 1: hipEvent timing around each kernel of kernelInvoke */
    EXC_OPT_COOP_MAX = 5    /* tpe 0: a list whose needed segments form a run of at most this many
                               is decoded wave-cooperatively, longer runs use random access */
};
int exc_ctx_set_option(exc_ctx *ctx, int key, int64_t value);
/* Times (ms) of the last kernelInvoke: probe kernel and the whole invoke. */
int exc_ctx_last_times(exc_ctx *ctx, float *probe_ms, float *invoke_ms);

/* ---- index residency (replaces as_load_atind + the whole-.ind2 cudaMemcpy,
 *      GC/LRC/LRC_Intersection/LRC_Intersection.cu:50-74,94-106)
 * A ParaPFD index of 128-docID blocks with medians also gets, per block, its
 * values at the other eighth boundaries for the sparse probe: its lists are
 * decoded on the device once, 2^28 docIDs at a time, through ctx's memory;
 * when ctx's budget leaves no room the index is uploaded without them. */
int exc_index_upload(exc_ctx *ctx, const exc_index_host *h, exc_index **out);
int exc_index_load_files(exc_ctx *ctx, int codec, uint32_t seg_size, const char *dir,
                         const char *name, exc_index **out);
void exc_index_free(exc_index *ix);
uint64_t exc_index_device_bytes(const exc_index *ix);
/* Number of lists and, when out is given (cap >= that number), each list's
 * length: the .ind1 m_urlcount column the reference's host side keeps to plan
 * batches (GC/LRC/LRC_Intersection/LRC_Intersection.cu:126-246). */
int exc_index_lengths(const exc_index *ix, uint32_t *out, uint32_t cap, uint32_t *nterms);

/* ---- batch stages, one call per reference stage
 *      (GC/LRC/LRC_Intersection/LRC_Intersection.cu:126,249,256,348,354).
 * CPUPreprocess parses the .query stream {u32 tnum; u32 tno[tnum]}... from
 * *ptr_in, stops once the summed shortest-list length exceeds threshold (as the
 * reference does), sorts each query's terms by ascending length and plans the
 * work units.  tnum is the full 32-bit word (the reference reads its low byte). */
int exc_CPUPreprocess(exc_ctx *ctx, exc_index *ix, const uint8_t **ptr_in,
                      const uint8_t *ptr_end, uint64_t threshold, exc_batchInfo *bi);
int exc_htodTransfer(exc_ctx *ctx, const exc_batchInfo *bi);
int exc_kernelInvoke(exc_ctx *ctx, const exc_batchInfo *bi);
/* ucntResult[q] = inclusive cumulative result count (the reference's h_ucntResult);
 * result = per-query results, ascending, concatenated in query order. */
int exc_dtohTransfer(exc_ctx *ctx, const exc_batchInfo *bi, uint32_t *ucntResult,
                     uint32_t *result, uint64_t result_cap);
int exc_verify(const exc_batchInfo *bi, const uint32_t *ucntResult, const uint32_t *result,
               exc_checkSum *cs);

/* All stages in one call for an in-memory query list (terms of query q are
 * qterms[qoff[q] .. qoff[q+1])). */
int exc_intersect(exc_ctx *ctx, exc_index *ix, const uint32_t *qterms, const uint64_t *qoff,
                  uint32_t nq, uint32_t *ucntResult, uint32_t *result, uint64_t result_cap,
                  uint64_t *n_results);
/* exc_intersect for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Same arguments and results as exc_intersect;
 * nothing is allocated on the device or the host: the batch is planned on the
 * device from the query list, inside a fixed 64 MiB scratch area that the
 * library holds on every GPU it runs on until the process ends (about 8 B per
 * shortest-list docID + 40 B per 512 of them + 110 B per query term; at most
 * 64 terms per query).  A larger batch returns EXC_ERR_CAPACITY before any
 * work is queued.  Calls that overlap on one GPU (several host threads or
 * contexts on the same device) race on the scratch area: serialise them. */
int exc_intersect_noalloc(exc_ctx *ctx, exc_index *ix, const uint32_t *qterms, const uint64_t *qoff,
                          uint32_t nq, uint32_t *ucntResult, uint32_t *result, uint64_t result_cap,
                          uint64_t *n_results);
/* Plan + upload a query list without running it (replay with exc_kernelInvoke). */
int exc_plan(exc_ctx *ctx, exc_index *ix, const uint32_t *qterms, const uint64_t *qoff,
             uint32_t nq, exc_batchInfo *bi);

/* ---- autotuner.  No reference counterpart: the reference hard-codes one
 *      launch shape per program (256 threads per shortest-list segment,
 *      GC/LRC/LRC_Intersection/LRC_Intersection.cu:256-283).
 * Queries fall into list-length-ratio classes by |second-shortest| /
 * |shortest|: class c < 4 for ratios in [10^c, 10^(c+1)), class 4 for ratios
 * >= 10^4, class 5 for single-term queries.  For every class present in the
 * sample, exc_autotune times the class's queries under each candidate of
 * (threads per short-list element as the cooperative-run limit, search mode)
 * and keeps the fastest; then it times the whole sample with the persistent
 * work queue on and off.  The result is installed in ctx (later plans use it,
 * EXC_OPT_PERSISTENT is set) and returned. */
#define EXC_RATIO_CLASSES 6
typedef struct {
    uint32_t queries[EXC_RATIO_CLASSES];  /* sample queries of the class                    */
    uint32_t coop_max[EXC_RATIO_CLASSES]; /* 0 = one lane per element ... 64 = wave-cooperative */
    uint32_t search[EXC_RATIO_CLASSES];   /* EXC_OPT_SEARCH value                            */
    float ms[EXC_RATIO_CLASSES];          /* class probe time with the choice               */
    float ms_default[EXC_RATIO_CLASSES];  /* ... with the context's options before tuning   */
    uint32_t persistent;                  /* chosen launch mode                              */
    float batch_ms[2];                    /* whole-sample probe time, non-persistent / persistent */
} exc_tune_table;
int exc_autotune(exc_ctx *ctx, exc_index *ix, const uint32_t *qterms, const uint64_t *qoff, uint32_t nq,
                 uint32_t reps, exc_tune_table *out);
/* Install a table (NULL: back to the plain options). */
int exc_ctx_set_tuning(exc_ctx *ctx, const exc_tune_table *table);

/* ---- routing one batch over candidate layouts of the same lists.  No
 *      reference counterpart: the reference fixes the hash-bucket width
 *      (UNI_LEN, src/gpu/Uncompressed/src/cuda_hashSegment_new/cuda_hashSegment.cu:42)
 *      and the segment size (SegmentSize, GC/LRC/LRC_Compression/main.cpp:40)
 *      at compile time, one program per choice.
 * The candidates are indexes of the same lists (same term count and list
 * lengths) under different layouts: segment sizes, bucket widths, codecs.  A
 * router sends every query to the candidate chosen for its ratio class (the
 * classes of exc_autotune).  Each candidate's sub-batch runs on a stream of its
 * own; all are queued before any is waited for, and the routed batch is timed
 * as one run, from a marker before the first kernel to one after the last.
 * With EXC_OPT_GRAPH the whole routed run is one captured hipGraph.  Results
 * come back in the caller's query order, as exc_intersect's.
 * exc_router_autotune tunes every candidate's knobs on the sample
 * (exc_autotune), routes each class to the candidate with the fastest class
 * time, then times the whole sample routed (both launch modes) and the whole
 * sample on each candidate alone, the same way; when one candidate alone is
 * faster than the routed run, every class goes to it.  A new router routes
 * every class to candidate 0. */
#define EXC_ROUTER_MAX_CANDIDATES 8
typedef struct exc_router exc_router;
typedef struct {
    uint32_t ncands;
    uint32_t route[EXC_RATIO_CLASSES];                            /* candidate of each ratio class     */
    float class_ms[EXC_ROUTER_MAX_CANDIDATES][EXC_RATIO_CLASSES]; /* class probe time, knobs tuned      */
    float single_ms[EXC_ROUTER_MAX_CANDIDATES];                   /* whole sample on one candidate      */
    float routed_ms;                                              /* whole sample routed by class_ms    */
    uint32_t persistent;                                          /* launch mode of the routed kernels  */
} exc_route_table;
int exc_router_create(exc_ctx *ctx, exc_index *const *cands, uint32_t ncands, exc_router **out);
int exc_router_autotune(exc_router *r, const uint32_t *qterms, const uint64_t *qoff, uint32_t nq,
                        uint32_t reps, exc_route_table *out);
/* route[c] = candidate of ratio class c, for all EXC_RATIO_CLASSES classes */
int exc_router_set_routes(exc_router *r, const uint32_t *route);
/* EXC_OPT_GRAPH is the router's own; every other option goes to each candidate's context */
int exc_router_set_option(exc_router *r, int key, int64_t value);
/* plan + upload (the stages of exc_plan), run the planned batch (ms = the
 * run's device time, may be NULL), fetch its results */
int exc_router_plan(exc_router *r, const uint32_t *qterms, const uint64_t *qoff, uint32_t nq);
int exc_router_invoke(exc_router *r, float *ms);
int exc_router_fetch(exc_router *r, uint32_t *ucntResult, uint32_t *result, uint64_t result_cap,
                     uint64_t *n_results);
int exc_router_intersect(exc_router *r, const uint32_t *qterms, const uint64_t *qoff, uint32_t nq,
                         uint32_t *ucntResult, uint32_t *result, uint64_t result_cap, uint64_t *n_results);
void exc_router_destroy(exc_router *r);

/* ---- decompression programs (GC/<codec>/<codec>_Decompression):
 * decode whole lists into out (concatenated in the given order). */
int exc_decompress(exc_ctx *ctx, exc_index *ix, const uint32_t *terms, uint32_t nterms,
                   uint32_t *out, uint64_t out_cap, uint64_t *n_out);
/* The same into device memory d_out (HIP device pointer of out_cap docIDs);
 * with EXC_OPT_TIMING the decode kernel's time is reported by
 * exc_ctx_last_times as probe_ms. */
int exc_decompress_device(exc_ctx *ctx, exc_index *ix, const uint32_t *terms, uint32_t nterms,
                          uint32_t *d_out, uint64_t out_cap, uint64_t *n_out);

/* ---- learned position models.  No reference counterpart: the reference's
 *      precursor is the single linear-regression window of
 *      src/gpu/Uncompressed/src/cuda_runRegressionAlgorithm/cuda_linearRegression.cu:792-873
 *      (and the list-level fit of .regression_info_list), used there without
 *      a fallback.  Every list longer than min_len gets a model trained on the
 *      GPU from the decoder's (docID, position) pairs; its recorded max error
 *      bounds the window of the search (EXC_OPT_SEARCH = 2), whose exact
 *      fallback keeps results identical to plain binary search. */
/* EXC_MODEL_PLA: 256 linear pieces at equal docID steps (knot p = the count
 * of elements below x0 + p / xs); EXC_MODEL_MLP: 1-32-1 ReLU network;
 * EXC_MODEL_PLA_Q: 256 linear pieces at equal position steps (knot j at
 * position floor(j (n - 1) / 256), its docID as a float), found by binary
 * search over the knots -- its error stays within about n / 256 positions
 * (plus the float spacing of the docIDs) whatever their distribution, where
 * a clustered list defeats equal docID steps. */
enum { EXC_MODEL_PLA = 1, EXC_MODEL_MLP = 2, EXC_MODEL_PLA_Q = 3 };
typedef struct { uint32_t nmodels; float max_err; float mean_err; float train_ms; } exc_model_stats;
typedef struct {
    int32_t kind;       /* EXC_MODEL_PLA (257 knots) | EXC_MODEL_MLP (w1[32] b1[32] w2[32] b2) |
                           EXC_MODEL_PLA_Q (x[257] docIDs, y[257] positions, s[256] piece slopes) */
    uint32_t n;         /* list length */
    float err;          /* recorded max |prediction - position| */
    float x0, xs;       /* input offset and scale (PLA, MLP) */
    uint32_t nparams;
    float params[770];
} exc_model_host;
int exc_index_train_models(exc_ctx *ctx, exc_index *ix, int kind, uint32_t min_len, uint32_t epochs,
                           uint32_t samples, float lr, uint64_t seed, exc_model_stats *stats);
int exc_index_model_info(exc_ctx *ctx, exc_index *ix, uint32_t term, exc_model_host *out);
/* One full-batch MLP gradient (the training kernel with zero epochs) on the
 * sorted list docs[0..n): grad[97] of the mean squared position error at params
 * over `samples` strided pairs; loss gets the mean squared error. */
int exc_mlp_gradient(exc_ctx *ctx, const float *params, const uint32_t *docs, uint32_t n, uint32_t samples,
                     float *grad, float *loss);

/* ---- serving: continuous batching over a resident index.  No reference
 *      counterpart (the reference batches a static query file by a
 *      shortest-list threshold, GC/LRC/LRC_Intersection/LRC_Intersection.cu:449-540).
 * Queries arrive at any time; a request loop admits them into free batch slots
 * (one HIP stream each) while the slots' device memory stays within
 * mem_budget, so batches overlap on the device; every query gets its own
 * results and latency.  nslots <= 16. */
typedef struct exc_server exc_server;
typedef struct { uint64_t batches, queries, max_inflight_batches, overlapped_launches; } exc_server_info;
/* flags: EXC_SERVER_COUNTS -- return per-query counts and latencies only (the
 * docIDs stay in device memory; for callers that rank or page on the device). */
#define EXC_SERVER_COUNTS 1u
int exc_server_create(exc_ctx *ctx, exc_index *ix, uint32_t nslots, uint32_t max_batch, uint64_t mem_budget,
                      uint32_t flags, exc_server **out);
int exc_server_submit(exc_server *sv, const uint32_t *terms, uint32_t nterms, uint64_t *ticket);
int exc_server_wait(exc_server *sv, uint64_t ticket, uint32_t *results, uint64_t cap, uint64_t *n,
                    double *latency_ms);
/* Submit queries at arrival_s[q] seconds from now (the caller's thread paces
 * them) and wait for all: per-query latency (ms from the scheduled arrival),
 * result counts and concatenated results. */
int exc_server_replay(exc_server *sv, const uint32_t *qterms, const uint64_t *qoff, const double *arrival_s,
                      uint32_t nq, double *latency_ms, uint32_t *counts, uint32_t *results, uint64_t res_cap);
int exc_server_stats(exc_server *sv, exc_server_info *st);
void exc_server_destroy(exc_server *sv);

/* ---- index construction on the GPU (replaces the reference's per-list
 *      generator loops with synchronous copies, GC/SegLRC/SegLRC_Generator/
 *      SegLRC_Generator.cu:94-331, GC/LRC/LRC_Generator/LRC_Generator.cu:144-188,
 *      GC/HS_SegLRC/HS_SegLRC_Generator/HS_SegLRC_Generator.cu:95-288,503-594,
 *      and the CPU packing of GC/LRC/LRC_Compression/main.cpp:163-199,
 *      GC/ParaPFD/ParaPFD_Compression/main.cpp:151-189 and
 *      GC/HS_SegLRC/HS_SegLRC_Compression/main.cpp:149-178, for indexes built
 *      where they are served).
 * exc_synth_lists_device: seeded synthetic lists, term t a Bernoulli(df[t] /
 *   universe) sample of [0, universe) (so its length is df[t] on average);
 *   list_off[0..nterms] receives the offsets; with d_docs (device memory of cap
 *   docIDs) the lists are written there, with d_docs NULL only counted.
 * exc_build_index_device: an index of codec EXC_SEGLRC, EXC_LRC, EXC_PARAPFD
 *   or EXC_HS_SEGLRC (frac: the ParaPFor exception fraction, ParaPFD only;
 *   hs_target and hs_shift: the HS bucket rule as for exc_build_index) built
 *   from sorted lists already in device memory (list t = d_docs[list_off[t] ..
 *   list_off[t+1]), strictly increasing, < 2^31); block-level arrays and HS
 *   tables byte-identical to exc_build_index (the LRC list fit too; SegLRC's
 *   list-level fit sums in another order).
 * exc_synth_index_device: both, in batches of about batch_postings docIDs, so
 *   that the raw lists never exist whole; list_off_out (nterms + 1) optional.
 * exc_index_export: a SegLRC, LRC, ParaPFD or HS_SegLRC index's arrays back on the host
 *   in the reference layout (for exc_index_save_files or comparison). */
int exc_synth_lists_device(exc_ctx *ctx, uint64_t seed, uint32_t universe, const uint32_t *df, uint32_t nterms,
                           uint64_t *list_off, uint32_t *d_docs, uint64_t cap);
int exc_build_index_device(exc_ctx *ctx, int codec, uint32_t seg, float frac, uint32_t hs_target, uint32_t hs_shift,
                           const uint32_t *d_docs, const uint64_t *list_off, uint32_t nterms, exc_index **out);
int exc_synth_index_device(exc_ctx *ctx, int codec, uint32_t seg, float frac, uint32_t hs_target, uint32_t hs_shift,
                           uint64_t seed, uint32_t universe, const uint32_t *df, uint32_t nterms,
                           uint64_t batch_postings, uint64_t *list_off_out, exc_index **out);
int exc_index_export(exc_ctx *ctx, exc_index *ix, exc_built **out);

/* ---- index builder (replaces the *_Generator and *_Compression programs,
 *      GC/LRC/LRC_Generator/LRC_Generator.cu:93-287, GC/LRC/LRC_Compression/main.cpp:86-238
 *      and their SegLRC / HS_SegLRC / ParaPFD counterparts).
 * docs: concatenated sorted distinct docID lists, list t = docs[list_off[t] .. list_off[t+1]).
 * frac: ParaPFor / NewPFor exception fraction (FRAC); hs_target: expected bucket size (256);
 * hs_shift: 0 = adaptive shift, else a fixed shift for every list.  Adaptive
 * shifts follow each program's own rule: HS_SegLRC bits(max) - bits(ceil(n /
 * hs_target)) (HS_SegLRC_Generator.cu:503-594), RAW_HS ceil(log2 max) -
 * ceil(log2(n / hs_target)) in double logs and a one-bucket table for lists of
 * at most hs_target docIDs (cuda_hashSegment.cu:872-905).  RAW_HS files:
 * <name>.ind1, <name>.ind2, <name>_HashSeg_offset, <name>_HashSeg_info. */
int exc_build_index(int codec, uint32_t seg_size, float frac, uint32_t hs_target,
                    uint32_t hs_shift, const uint32_t *docs, const uint64_t *list_off,
                    uint32_t nterms, int nthreads, exc_built **out);
int exc_built_view(const exc_built *b, exc_index_host *view);
void exc_built_free(exc_built *b);
/* Write / read the reference's file set <dir>/<name>.{ind1,ind2,segNum,...}. */
int exc_index_save_files(const exc_index_host *h, const char *dir, const char *name);

/* ---- synthetic corpus (stands in for the reference's GP/ preprocessing tools):
 * list t = df[t] distinct docIDs drawn uniformly from [0, universe), sorted,
 * written to docs[list_off[t] ..].  Deterministic in (seed, t). */
int exc_synth_lists(uint64_t seed, uint32_t universe, const uint32_t *df, uint32_t nterms,
                    int nthreads, uint32_t *docs, const uint64_t *list_off);

#ifdef __cplusplus
}
#endif
#endif
