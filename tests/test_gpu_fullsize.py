"""GPU parity at the full C2 size: 100,000 Zipf-distributed lists over 2^25
docIDs (369 M postings), the 1,000-query C2 batch of 2-5 terms.

For SegLRC-128, LRC-256 and ParaPFD-128 (FRAC 0.1) the device's cumulative
counts and results must equal the oracle's exactly (its C restatement of
decode + intersect, multithreaded, over the same encoded arrays), in both
launch modes; the uncompressed lists must give the same output, and so must
the C3 hash buckets (shift 13 and the adaptive width) for a share of the
ratio queries.  The SegLRC
results are further checked against the raw lists themselves: each query's
results ascend strictly and every result is in every list of its query.
Whole-list decoding of all 100,000 lists must return them unchanged for every
codec, NewPFD included, the device encoder must produce the host builder's
segment arrays byte for byte, the index written to the reference's files
must answer the same after loading, and so must the learned window search
(PLA and MLP models on every list longer than 65,536) and the
continuous-batching server."""
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


def _threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


@pytest.fixture(scope="module")
def c2():
    df, docs, off = W.c2_corpus()
    assert len(docs) > 369_000_000
    queries = W.c2_queries(df)
    built = ea.build_index(docs, off, ea.SEGLRC, seg=128)
    cum, res = oracle.eval_queries(_oracle_index(built), queries, nthreads=_threads())
    return docs, off, queries, built, np.asarray(cum, np.int64), res, df


def _gpu_both_modes(ctx, built, queries):
    dix = ctx.upload(built)
    out = []
    try:
        for persistent in (0, 1):
            ctx.set_option(ea.OPT_PERSISTENT, persistent)
            cum, res = ctx.intersect(dix, queries)
            out.append((np.asarray(cum, np.int64), res))
    finally:
        ctx.set_option(ea.OPT_PERSISTENT, 1)
        dix.close()
    return out


def test_c2_full_size_seglrc(ctx, c2):
    docs, off, queries, built, ocum, ores, _ = c2
    for mode, (cum, res) in enumerate(_gpu_both_modes(ctx, built, queries)):
        assert np.array_equal(cum, ocum), mode
        assert np.array_equal(res, ores), mode
    # properties against the raw lists
    start = np.concatenate([[0], ocum[:-1]])
    for q, terms in enumerate(queries):
        r = ores[start[q]:ocum[q]]
        assert np.all(np.diff(r.astype(np.int64)) > 0), q
        for t in terms:
            lst = docs[off[t]:off[t + 1]]
            pos = np.minimum(np.searchsorted(lst, r), max(len(lst) - 1, 0))
            assert len(r) == 0 or np.array_equal(lst[pos], r), (q, t)


@pytest.mark.parametrize("codec,seg,frac", [(ea.LRC, 256, 0.0), (ea.PARAPFD, 128, 0.1), (ea.RAW, 128, 0.0)])
def test_c2_full_size_other_codecs(ctx, c2, codec, seg, frac):
    docs, off, queries, _, scum, sres, _ = c2
    built = ea.build_index(docs, off, codec, seg=seg, frac=frac)
    if codec == ea.RAW:
        ocum, ores = scum, sres  # the uncompressed lists answer like any encoding of them
    else:
        ocum, ores = oracle.eval_queries(_oracle_index(built), queries, nthreads=_threads())
        ocum = np.asarray(ocum, np.int64)
        assert np.array_equal(ocum, scum) and np.array_equal(ores, sres)
    for mode, (cum, res) in enumerate(_gpu_both_modes(ctx, built, queries)):
        assert np.array_equal(cum, ocum), mode
        assert np.array_equal(res, ores), mode


@pytest.mark.parametrize("shift", [13, 0])
def test_c3_full_size_hash_buckets(ctx, c2, shift):
    """C3: the full-size lists in hash buckets of shift 13 (the top 12 bits of
    2^25) or of the adaptive width (about 256 docIDs per bucket, per list), all
    10,000 ratio queries (long:short 10^U(1,4), seed 3): equal to the oracle
    over the same buckets in both launch modes, and as the configuration runs
    them -- one persistent work-queue launch replayed from a hipGraph, knobs
    autotuned per ratio class -- on two replays."""
    docs, off, _, _, _, _, df = c2
    queries = W.ratio_queries(df, 10_000, 3)
    built = ea.build_index(docs, off, ea.HS_SEGLRC, seg=128, hs_shift=shift)
    ocum, ores = oracle.eval_queries(_oracle_index(built), queries, nthreads=_threads())
    ocum = np.asarray(ocum, np.int64)
    dix = ctx.upload(built)
    if shift == 13:
        # records only for buckets that hold docIDs: 4.47 GB on the device,
        # against 15.5 GB with a record for every bucket
        assert dix.device_bytes() < 5.0e9, dix.device_bytes()
    try:
        for persistent in (0, 1):
            ctx.set_option(ea.OPT_PERSISTENT, persistent)
            cum, res = ctx.intersect(dix, queries)
            assert np.array_equal(np.asarray(cum, np.int64), ocum), persistent
            assert np.array_equal(res, ores), persistent
        ctx.autotune(dix, W.ratio_queries(df, 2000, 33), reps=1)
        ctx.set_option(ea.OPT_PERSISTENT, 1)
        ctx.set_option(ea.OPT_GRAPH, 1)
        ctx.plan(dix, queries)
        bi = ctx._bi
        for rep in range(2):
            ctx.kernelInvoke(bi)
            cum, res = ctx.dtohTransfer(bi)
            assert np.array_equal(np.asarray(cum, np.int64), ocum), rep
            assert np.array_equal(res, ores), rep
    finally:
        ctx.set_option(ea.OPT_GRAPH, 0)
        ctx.set_tuning(None)
        ctx.set_option(ea.OPT_SEARCH, 1)
        ctx.set_option(ea.OPT_PERSISTENT, 1)
        dix.close()


def test_c3_full_size_routed(ctx, c2):
    """C3 with the bucket width chosen per ratio class: the router over the
    full-size lists in buckets of shift 13, of the adaptive width and in
    SegLRC-128 segments, routes tuned on a 2,000-query sample, all 10,000
    ratio queries run as one captured graph of the routed sub-batches (two
    replays), every docID equal to the oracle's."""
    docs, off, _, built, _, _, df = c2
    queries = W.ratio_queries(df, 10_000, 3)
    ocum, ores = oracle.eval_queries(_oracle_index(built), queries, nthreads=_threads())
    ocum = np.asarray(ocum, np.int64)
    hs = [ea.build_index(docs, off, ea.HS_SEGLRC, seg=128, hs_shift=sh) for sh in (13, 0)]
    dixs = [ctx.upload(b) for b in hs] + [ctx.upload(built)]
    router = ea.Router(ctx, dixs)
    try:
        t = router.autotune(W.ratio_queries(df, 2000, 33), reps=1)
        print("routes", t["route"], "routed_ms", t["routed_ms"], "single_ms", t["single_ms"], flush=True)
        router.set_option(ea.OPT_GRAPH, 1)
        router.plan(queries)
        for rep in range(2):
            router.invoke()
            cum, res = router.fetch()
            assert np.array_equal(np.asarray(cum, np.int64), ocum), rep
            assert np.array_equal(res, ores), rep
    finally:
        router.close()
        for d in dixs:
            d.close()
        for b in hs:
            b.close()


@pytest.mark.parametrize("shift", [13, 0])
def test_c3_full_size_uncompressed_hash_buckets(ctx, c2, shift):
    """C3 on the uncompressed hash-segment layout (the docIDs themselves plus
    [shift][start_0 .. start_nb] tables: shift 13 = the top 12 bits of 2^25, or
    the reference generator's own adaptive shift): all 10,000 ratio queries
    equal to the oracle in both launch modes, binary and interpolation search
    inside the buckets."""
    docs, off, _, built, _, _, df = c2
    queries = W.ratio_queries(df, 10_000, 3)
    ocum, ores = oracle.eval_queries(_oracle_index(built), queries, nthreads=_threads())
    ocum = np.asarray(ocum, np.int64)
    rb = ea.build_index(docs, off, ea.RAW_HS, hs_shift=shift)
    dix = ctx.upload(rb)
    try:
        for persistent in (0, 1):
            for search in (0, 1):
                ctx.set_option(ea.OPT_PERSISTENT, persistent)
                ctx.set_option(ea.OPT_SEARCH, search)
                cum, res = ctx.intersect(dix, queries)
                assert np.array_equal(np.asarray(cum, np.int64), ocum), (persistent, search)
                assert np.array_equal(res, ores), (persistent, search)
    finally:
        ctx.set_option(ea.OPT_SEARCH, 1)
        ctx.set_option(ea.OPT_PERSISTENT, 1)
        dix.close()
        rb.close()


def test_c4_full_query_set(ctx, c2):
    """C4's own query set: 10,000 queries of 2-4 terms (seed 44) over the
    full-size SegLRC index, with PLA-256 and then MLP 1-32-1 models on every
    list longer than 65,536; binary, list-fit and learned window search each
    equal to the oracle (zero mismatches)."""
    docs, off, _, built, _, _, df = c2
    queries = W.df_weighted_queries(df, 10_000, 2, 4, 44)
    ocum, ores = oracle.eval_queries(_oracle_index(built), queries, nthreads=_threads())
    ocum = np.asarray(ocum, np.int64)
    dix = ctx.upload(built)
    try:
        for kind in (ea.MODEL_PLA, ea.MODEL_MLP):
            ctx.train_models(dix, kind, min_len=65536, epochs=200, samples=4096, lr=1e-3)
            for search in (ea.SEARCH_BINARY, 1, ea.SEARCH_LEARNED):
                ctx.set_option(ea.OPT_SEARCH, search)
                cum, res = ctx.intersect(dix, queries)
                assert np.array_equal(np.asarray(cum, np.int64), ocum), (kind, search)
                assert np.array_equal(res, ores), (kind, search)
    finally:
        ctx.set_option(ea.OPT_SEARCH, 1)
        dix.close()


@pytest.mark.parametrize("codec,seg,frac", [(ea.SEGLRC, 128, 0.0), (ea.LRC, 256, 0.0), (ea.PARAPFD, 128, 0.1),
                                            (ea.NEWPFD, 256, 0.1)])
def test_c2_full_size_decode(ctx, c2, codec, seg, frac):
    """Every one of the 100,000 full-size lists decoded on the device in one
    call equals the list it was built from."""
    docs, off, _, built0, _, _, _ = c2
    built = built0 if codec == ea.SEGLRC else ea.build_index(docs, off, codec, seg=seg, frac=frac)
    dix = ctx.upload(built)
    try:
        out = ctx.decompress(dix, np.arange(len(off) - 1, dtype=np.uint32))
    finally:
        dix.close()
    assert np.array_equal(out, docs)


@pytest.mark.parametrize("codec,seg,frac", [(ea.SEGLRC, 128, 0.0), (ea.LRC, 256, 0.0), (ea.PARAPFD, 128, 0.1),
                                            (ea.HS_SEGLRC, 128, 0.0)])
def test_c2_full_size_device_builder(ctx, c2, codec, seg, frac):
    """The device encoders over the full-size lists (HS_SegLRC with the
    adaptive bucket width): the index's arrays in the reference layout
    byte-identical to the host builder's, and the C2 batch answered on the
    device-built index exactly as the oracle answers it."""
    docs, off, queries, built, ocum, ores, _ = c2
    host = built if codec == ea.SEGLRC else ea.build_index(docs, off, codec, seg=seg, frac=frac)
    names = {ea.SEGLRC: ("seg_info", "seg_base", "seg_head"), ea.LRC: ("lrc_info", "seg_head"),
             ea.PARAPFD: ("seg_median", "seg_head"),
             ea.HS_SEGLRC: ("seg_info", "seg_base", "hash_offset", "hash_info")}[codec]
    d = ctx.device_alloc(len(docs) * 4)
    try:
        ctx.to_device(d, docs)
        dix = ctx.build_index_device(d, off, seg=seg, codec=codec, frac=frac)
        try:
            dev = ctx.export(dix)
            for name in ("ind1", "ind2", "seg_num", "seg_offset") + names:
                assert getattr(dev, name).tobytes() == getattr(host, name).tobytes(), name
            cum, res = ctx.intersect(dix, queries)
            assert np.array_equal(np.asarray(cum, np.int64), ocum) and np.array_equal(res, ores)
        finally:
            dix.close()
    finally:
        ctx.device_free(d)


def test_c2_full_size_from_files(ctx, c2, tmp_path):
    """The full-size SegLRC-128 index written in the reference's files and
    loaded back through the mapped, validated, staged loader answers the C2
    batch as the oracle does."""
    docs, off, queries, built, ocum, ores, _ = c2
    built.save(str(tmp_path), "c2")
    lix = ctx.load_files(ea.SEGLRC, 128, str(tmp_path), "c2")
    try:
        cum, res = ctx.intersect(lix, queries)
        assert np.array_equal(np.asarray(cum, np.int64), ocum) and np.array_equal(res, ores)
    finally:
        lix.close()


@pytest.mark.parametrize("kind", [ea.MODEL_PLA, ea.MODEL_MLP])
def test_c2_full_size_learned_search(ctx, c2, kind):
    """C4's learned window search at full size: a position model on every list
    longer than 65,536 (511 lists), the C2 batch answered through the models'
    windows exactly as the oracle answers it."""
    docs, off, queries, built, ocum, ores, _ = c2
    dix = ctx.upload(built)
    try:
        if kind == ea.MODEL_PLA:
            st = ctx.train_models(dix, kind, min_len=65536)
        else:
            st = ctx.train_models(dix, kind, min_len=65536, epochs=200, samples=4096, lr=1e-3)
        assert st["nmodels"] == int((np.diff(off.astype(np.int64)) > 65536).sum())
        ctx.set_option(ea.OPT_SEARCH, ea.SEARCH_LEARNED)
        cum, res = ctx.intersect(dix, queries)
        assert np.array_equal(np.asarray(cum, np.int64), ocum) and np.array_equal(res, ores)
    finally:
        ctx.set_option(ea.OPT_SEARCH, 1)
        dix.close()


def test_c2_full_size_served(ctx, c2):
    """The C2 batch served by the continuous-batching loop at full size (three
    slots, Poisson arrivals at 20 k queries/s): every query's count and
    results equal the oracle's."""
    docs, off, queries, built, ocum, ores, _ = c2
    dix = ctx.upload(built)
    try:
        arrival = np.cumsum(np.random.default_rng(5).exponential(1 / 20000.0, len(queries)))
        srv = ea.Server(ctx, dix, nslots=3, max_batch=256)
        try:
            lat, cnt, res = srv.replay(queries, arrival, int(ocum[-1]))
        finally:
            srv.close()
        assert np.array_equal(cnt.astype(np.int64), np.diff(np.concatenate([[0], ocum])))
        assert np.array_equal(res, ores) and np.all(lat > 0)
    finally:
        dix.close()
