/*
 * oracle.h -- CPU restatement of ExCLucene's compressed posting-list codecs
 * and exact k-term intersection (src/gpu in the reference).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the exclucene_amd package links,
 * loads or calls this code; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker.
 *
 * Reference files restated here (paths relative to the reference root):
 *   GI = src/gpu/Compressed/include, GC = src/gpu/Compressed/src
 *   GI/ParaPFor.cpp:31-134, GI/pack.h:14-24, GI/unpack.h:14-23, GI/util.h:13-27
 *   GC/LRC/LRC_Generator/LRC_Generator.cu:165-177,221-287
 *   GC/LRC/LRC_Generator/LRC_Generator_kernel.cu:55-91
 *   GC/LRC/LRC_Compression/main.cpp:163-199
 *   GC/SegLRC/SegLRC_Generator/SegLRC_Generator.cu:185-233,264-331
 *   GC/SegLRC/SegLRC_Generator/SegLRC_Generator_kernel.cu:57-103
 *   GC/HS_SegLRC/HS_SegLRC_Generator/HS_SegLRC_Generator.cu:243-288,503-594
 *   GC/HS_SegLRC/HS_SegLRC_Compression/main.cpp:149-178
 *   GC/ParaPFD/ParaPFD_Compression/main.cpp:151-189,235-311
 *   GI/NewPFor.cpp:54-133, GI/Simple16.h:13-258,
 *   GC/NewPFD/NewPFD_Compression/main.cpp:151-189,
 *   GC/NewPFD/NewPFD_Decompression/NewPFD_Decompression_kernel.cu:405-470
 *   GC/LRC/LRC_Intersection/LRC_Intersection.cu:126-215 (query semantics),
 *   :354-379 (checksums)
 *
 * Pinned fp32 reconstruction formulas (the reference leaves them to nvcc's
 * default FMA contraction; this restatement fixes them explicitly):
 *   LRC    : y(i) = (int32)((float)i * slope)            + icpt
 *   SegLRC : y(j) = (int32)fmaf((float)j, slope, icpt)    + base
 */
#ifndef EXC_ORACLE_H
#define EXC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_LRC = 0, ORC_SEGLRC = 1, ORC_HS = 2, ORC_PFD = 3, ORC_NPFD = 4 };

/* LRC per-list fit, 12 bytes (LRC_Intersection_kernel.cu:89-93). */
typedef struct { float fSlope; int32_t iIntercept; float fSearchRange; } orc_lrc_info;
/* SegLRC / HS per-segment (per-bucket) fit, 16 bytes (SegLRC_Generator_kernel.cu:37-42). */
typedef struct { float fSlope; float fIntercept; uint32_t nRangeLeft; uint32_t nRangeRight; } orc_seg_info;

/* ---- ParaPFor block codec (GI/ParaPFor.cpp) ---- */
uint32_t orc_pfor_best_b(const uint32_t *in, uint32_t n, float frac);
/* returns words written (descriptor included), 0 on an unrepresentable block */
uint32_t orc_pfor_encode(const uint32_t *in, uint32_t n, float frac, uint32_t *out);
/* returns words consumed */
uint32_t orc_pfor_decode(const uint32_t *in, uint32_t n, uint32_t *out);

/* ---- NewPFor block codec (GI/NewPFor.cpp) and Simple16 (GI/Simple16.h) ---- */
uint32_t orc_s16_encode(const uint32_t *in, uint32_t n, uint32_t *out);   /* words, 0 if a value >= 2^28 */
uint32_t orc_s16_decode(const uint32_t *in, uint32_t n, uint32_t *out);   /* words; writes <= n + 27 */
uint32_t orc_npfor_encode(const uint32_t *in, uint32_t n, float frac, uint32_t *out);
uint32_t orc_npfor_decode(const uint32_t *in, uint32_t n, uint32_t *out);
uint64_t orc_npfd_encode_list(const uint32_t *L, uint32_t n, uint32_t seg, float frac,
                              uint32_t *words, uint32_t *seg_head, uint32_t *seg_median,
                              uint32_t *seg_off);
int orc_npfd_decode_list(const uint32_t *words, uint32_t n, uint32_t seg,
                         const uint32_t *seg_head, const uint32_t *seg_off, uint32_t *out);

/* ---- LRC (one global fit per list, residual blocks of `seg`) ---- */
void orc_lrc_fit(const uint32_t *L, uint32_t n, orc_lrc_info *info);
uint64_t orc_lrc_encode_list(const uint32_t *L, uint32_t n, uint32_t seg, float frac,
                             orc_lrc_info *info, uint32_t *words,
                             uint32_t *seg_head, uint32_t *seg_off);
int orc_lrc_decode_list(const uint32_t *words, uint32_t n, uint32_t seg,
                        const orc_lrc_info *info, const uint32_t *seg_off, uint32_t *out);

/* ---- SegLRC (per-segment fit + per-segment base) ---- */
void orc_seg_fit(const uint32_t *L, uint32_t n, orc_seg_info *info);
uint64_t orc_seglrc_encode_list(const uint32_t *L, uint32_t n, uint32_t seg, float frac,
                                orc_seg_info *seg_fit, int32_t *base, uint32_t *words,
                                uint32_t *seg_head, uint32_t *seg_off, orc_seg_info *list_fit);
int orc_seglrc_decode_list(const uint32_t *words, uint32_t n, uint32_t seg,
                           const orc_seg_info *seg_fit, const int32_t *base,
                           const uint32_t *seg_off, uint32_t *out);

/* ---- HS_SegLRC (hash buckets on the top docID bits) ---- */
/* hash table words: [shift][nb][start_0 .. start_nb]; returns word count.
 * fixed_shift == 0 selects the reference's adaptive rule. */
uint32_t orc_hs_table(const uint32_t *L, uint32_t n, uint32_t bucket_target,
                      uint32_t fixed_shift, uint32_t *info);
/* uncompressed hash-segment table [shift][start_0 .. start_nb]
 * (cuda_hashSegment.cu:815-947); returns word count (<= (max >> shift) + 3). */
uint32_t orc_gu_hashseg_table(const uint32_t *L, uint32_t n, uint32_t target, uint32_t fixed_shift,
                              uint32_t *info);
uint64_t orc_hs_encode_list(const uint32_t *L, uint32_t n, const uint32_t *info, float frac,
                            orc_seg_info *fit, int32_t *base, uint32_t *words, uint32_t *seg_off);
int orc_hs_decode_list(const uint32_t *words, uint32_t n, const uint32_t *info,
                       const orc_seg_info *fit, const int32_t *base,
                       const uint32_t *seg_off, uint32_t *out);

/* ---- ParaPFD (d-gaps, ParaPFor blocks, segHead/segMedian) ---- */
uint64_t orc_pfd_encode_list(const uint32_t *L, uint32_t n, uint32_t seg, float frac,
                             uint32_t *words, uint32_t *seg_head, uint32_t *seg_median,
                             uint32_t *seg_off);
int orc_pfd_decode_list(const uint32_t *words, uint32_t n, uint32_t seg,
                        const uint32_t *seg_head, const uint32_t *seg_off, uint32_t *out);

/* ---- whole index + queries ---- */
typedef struct {
    int32_t codec;              /* ORC_* */
    uint32_t seg;               /* segment size (LRC/SegLRC/PFD) */
    uint32_t nterms;
    uint32_t pad_;
    uint64_t total_words;       /* words in the compressed ind2 */
    const uint32_t *ucnt;       /* per term: list length (ind1.m_urlcount) */
    const uint64_t *uwlist;     /* per term: word offset of the list in words */
    const uint32_t *segnum;     /* per term: index of its first segment / bucket */
    const uint32_t *words;      /* compressed ind2 */
    const uint32_t *seg_off;    /* per segment/bucket: word offset inside its list */
    const uint32_t *seg_head;   /* per segment: first docID (LRC/SegLRC/PFD) */
    const uint32_t *seg_median; /* per segment (PFD) */
    const orc_lrc_info *lrc;    /* per term (LRC) */
    const orc_seg_info *fit;    /* per segment / bucket (SegLRC/HS) */
    const int32_t *base;        /* per segment / bucket (SegLRC/HS) */
    const uint32_t *hash_off;   /* per term offset into hash_info (HS) */
    const uint32_t *hash_info;  /* HS tables */
} orc_index;

int orc_decode_term(const orc_index *ix, uint32_t term, uint32_t *out);

/* Reference checksum triple (LRC_Intersection.cu:354-379). */
void orc_checksum(const uint32_t *cum_counts, const uint32_t *results, uint32_t nq,
                  uint64_t cs[3]);

/* Evaluate nq queries (terms of query q are qterms[qoff[q] .. qoff[q+1])).
 * Writes inclusive cumulative result counts (the reference's h_ucntResult) and
 * the concatenated per-query ascending results (capacity res_cap words).
 * Block-skipping CPU decode-and-intersect, `nthreads` POSIX threads with a
 * static partition of queries.  Per-query wall times (ms) go to q_ms if non-NULL.
 * Returns 0, or -1 on capacity overflow / bad input. */
int orc_eval_queries(const orc_index *ix, const uint32_t *qterms, const uint64_t *qoff,
                     uint32_t nq, int nthreads, uint32_t *cum_counts, uint32_t *results,
                     uint64_t res_cap, double *q_ms);

/* Same queries by full decode of every list + merge intersection (the slow judge). */
int orc_eval_queries_full(const orc_index *ix, const uint32_t *qterms, const uint64_t *qoff,
                          uint32_t nq, uint32_t *cum_counts, uint32_t *results, uint64_t res_cap);

/* Sorted-set intersection of two ascending uint32 arrays; returns count. */
uint64_t orc_intersect2(const uint32_t *a, uint64_t na, const uint32_t *b, uint64_t nb, uint32_t *out);

/* Algorithmic bytes: compressed bytes of the shortest list and
 * its metadata + distinct blocks of each longer list containing a probe target
 * + 4 B per result. Returns total over the queries; full_bytes gets the
 * decode-everything figure. */
uint64_t orc_query_bytes(const orc_index *ix, const uint32_t *qterms, const uint64_t *qoff,
                         uint32_t nq, uint64_t *full_bytes);

/* ---- learned position models (no reference counterpart; checker for the
 * GPU-trained models).  Same fp32 operations, in the same order, as the
 * device inference, so recorded max errors must agree exactly. */
void orc_pla_fit(const uint32_t *L, uint32_t n, float *x0, float *xs, float *y /* 257 */);
float orc_pla_pred(const float *y, float x0, float xs, uint32_t v);
float orc_mlp_pred(const float *w /* 97 */, float x0, float xs, uint32_t n, uint32_t v);
void orc_plaq_fit(const uint32_t *L, uint32_t n, float *params /* 770 */);
float orc_plaq_pred(const float *params, uint32_t v);
float orc_model_max_err(int kind, const float *params, float x0, float xs, const uint32_t *L, uint32_t n);

#ifdef __cplusplus
}
#endif
#endif
