"""GPU parity at the full C2 size: 100,000 Zipf-distributed lists over 2^25
docIDs (369 M postings), SegLRC-128, the 1,000-query C2 batch of 2-5 terms.

The device's cumulative counts and results must equal the oracle's exactly
(its C restatement of decode + intersect, multithreaded, over the same
encoded arrays), for both launch modes, and must satisfy properties checked
against the raw lists themselves: each query's results ascend strictly and
every result is in every list of its query."""
import os

import numpy as np
import pytest

import exclucene_amd as ea
from exclucene_amd import workloads as W
import oracle

pytestmark = pytest.mark.gpu


def _oracle_index(built):
    return oracle.Index(codec=built.codec, seg=built.seg, ucnt=np.ascontiguousarray(built.ind1["m_urlcount"]),
                        uwlist=np.ascontiguousarray(built.ind1["m_off"] // 4), segnum=built.seg_num,
                        words=built.ind2, total_words=len(built.ind2), seg_off=built.seg_offset,
                        seg_head=built.seg_head, seg_median=built.seg_median, lrc=built.lrc_info,
                        fit=built.seg_info, base=built.seg_base, hash_off=built.hash_offset,
                        hash_info=built.hash_info)


def test_c2_full_size_matches_oracle(ctx):
    df, docs, off = W.c2_corpus()
    assert len(docs) > 369_000_000
    queries = W.c2_queries(df)
    built = ea.build_index(docs, off, ea.SEGLRC, seg=128)
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    ocum, ores = oracle.eval_queries(_oracle_index(built), queries, nthreads=threads)
    dix = ctx.upload(built)
    try:
        for persistent in (0, 1):
            ctx.set_option(ea.OPT_PERSISTENT, persistent)
            cum, res = ctx.intersect(dix, queries)
            assert np.array_equal(np.asarray(cum, np.int64), np.asarray(ocum, np.int64)), persistent
            assert np.array_equal(res, ores), persistent
    finally:
        ctx.set_option(ea.OPT_PERSISTENT, 1)
        dix.close()

    # properties against the raw lists
    cum = np.asarray(cum, np.int64)
    start = np.concatenate([[0], cum[:-1]])
    for q, terms in enumerate(queries):
        r = res[start[q]:cum[q]]
        assert np.all(np.diff(r.astype(np.int64)) > 0), q
        for t in terms:
            lst = docs[off[t]:off[t + 1]]
            pos = np.minimum(np.searchsorted(lst, r), max(len(lst) - 1, 0))
            assert len(r) == 0 or np.array_equal(lst[pos], r), (q, t)
