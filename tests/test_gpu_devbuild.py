"""GPU tests of index construction on the device: seeded synthetic lists, the
SegLRC, LRC, ParaPFD and HS_SegLRC encoders (the reference layout's arrays
byte-identical to the host builder's), the batched synthesize-and-encode path,
and exact queries on device-built indexes."""
import numpy as np
import pytest

import exclucene_amd as ea
import oracle

pytestmark = pytest.mark.gpu

U = 1 << 24
DF = np.array([0, 1, 2, 7, 100, 999, 4096, 4097, 20000, 70000, 300000, 524288, 600000, 2000000, 5000000, 16777216,
               3, 12345, 8191, 777777], np.uint32)


def _synth(ctx, seed, df=DF, universe=U):
    off = ctx.synth_lists_device(seed, universe, df)
    total = int(off[-1])
    d = ctx.device_alloc(max(total, 1) * 4)
    off2 = ctx.synth_lists_device(seed, universe, df, d, total)
    assert np.array_equal(off, off2)
    docs = ctx.from_device(d, total)
    return d, off, docs


def test_synth_lists_properties(ctx):
    d, off, docs = _synth(ctx, 11)
    try:
        for t, want in enumerate(DF):
            L = docs[int(off[t]):int(off[t + 1])]
            assert np.all(L < U)
            assert np.all(np.diff(L.astype(np.int64)) > 0), t
            p = want / U
            sd = np.sqrt(U * p * (1 - p))
            assert abs(len(L) - int(want)) <= 6 * sd + 6, (t, len(L), int(want))
        assert len(docs[int(off[15]):int(off[16])]) == U  # p = 1 keeps every docID
        d2, off2, docs2 = _synth(ctx, 11)
        assert np.array_equal(off, off2) and np.array_equal(docs, docs2)  # seeded
        ctx.device_free(d2)
        d3, off3, docs3 = _synth(ctx, 12)
        assert not np.array_equal(docs, docs3)
        ctx.device_free(d3)
    finally:
        ctx.device_free(d)


@pytest.mark.parametrize("seg", [128, 256])
def test_device_encoder_matches_host_builder(ctx, seg):
    d, off, docs = _synth(ctx, 21)
    try:
        host = ea.build_index(docs, off, ea.SEGLRC, seg=seg)
        dix = ctx.build_index_device(d, off, seg=seg)
        try:
            dev = ctx.export(dix)
            for name in ("ind1", "ind2", "seg_num", "seg_offset", "seg_head", "seg_info", "seg_base"):
                assert np.array_equal(getattr(dev, name), getattr(host, name)), name
            # the list-level fit sums in a different order: close, not identical
            a, b = dev.list_info, host.list_info
            big = host.ucnt > 1
            assert np.allclose(a["fSlope"][big], b["fSlope"][big], rtol=1e-5)
            assert np.all(np.abs(a["nRangeLeft"][big].astype(np.int64) - b["nRangeLeft"][big]) <= 2)
            # whole lists decode back
            out = ctx.decompress(ea.DeviceIndex(ctx, dix.handle, host), list(range(len(DF))))
            assert np.array_equal(out, docs)
        finally:
            dix.close()
    finally:
        ctx.device_free(d)


def test_device_built_index_answers_queries(ctx):
    d, off, docs = _synth(ctx, 31)
    try:
        dix = ctx.build_index_device(d, off, seg=128)
        lists = [docs[int(off[t]):int(off[t + 1])] for t in range(len(DF))]
        rng = np.random.default_rng(4)
        qs = [[int(x) for x in rng.choice(len(DF), int(rng.integers(1, 5)), replace=False)] for _ in range(300)]
        ocum, ores = oracle.eval_queries(oracle.build_index(lists, oracle.SEGLRC, seg=128), qs)
        try:
            for search in (0, 1):
                ctx.set_option(ea.OPT_SEARCH, search)
                cum, res = ctx.intersect(dix, qs)
                assert np.array_equal(cum, ocum) and np.array_equal(res, ores), search
        finally:
            ctx.set_option(ea.OPT_SEARCH, 1)
            dix.close()
    finally:
        ctx.device_free(d)


def test_batched_synth_index_equals_single_pass(ctx):
    d, off, docs = _synth(ctx, 41)
    try:
        one = ctx.build_index_device(d, off, seg=128)
        many, off2 = ctx.synth_index_device(41, U, DF, seg=128, batch_postings=1 << 20)
        try:
            assert np.array_equal(off, off2)
            a, b = ctx.export(one), ctx.export(many)
            for name in ("ind1", "ind2", "seg_num", "seg_offset", "seg_head", "seg_info", "seg_base", "list_info"):
                assert np.array_equal(getattr(a, name), getattr(b, name)), name
        finally:
            one.close()
            many.close()
    finally:
        ctx.device_free(d)


@pytest.mark.parametrize("codec,seg,frac", [(ea.LRC, 256, 0.0), (ea.LRC, 128, 0.0), (ea.PARAPFD, 128, 0.1),
                                            (ea.PARAPFD, 256, 0.1), (ea.PARAPFD, 128, 0.5), (ea.PARAPFD, 128, 0.0)])
def test_device_lrc_parapfd_encoders_match_host_builder(ctx, codec, seg, frac):
    """The LRC and ParaPFD encoders on the device: every array the reference's
    files hold (.ind1, .ind2, .segNum, .segOffset, .segHead, and LRC's
    .regression_info or ParaPFD's .segMedian) byte-identical to the host
    builder's, the index answering queries as the oracle does, and the
    batched synthesize-and-encode path equal to the one-pass build."""
    d, off, docs = _synth(ctx, 51)
    try:
        host = ea.build_index(docs, off, codec, seg=seg, frac=frac)
        dix = ctx.build_index_device(d, off, seg=seg, codec=codec, frac=frac)
        try:
            dev = ctx.export(dix)
            names = ["ind1", "ind2", "seg_num", "seg_offset", "seg_head"]
            names += ["lrc_info"] if codec == ea.LRC else ["seg_median"]
            for name in names:
                assert getattr(dev, name).tobytes() == getattr(host, name).tobytes(), name
            lists = [docs[int(off[t]):int(off[t + 1])] for t in range(len(DF))]
            rng = np.random.default_rng(5)
            qs = [[int(x) for x in rng.choice(len(DF), int(rng.integers(1, 5)), replace=False)] for _ in range(200)]
            ocum, ores = oracle.eval_queries(oracle.build_index(lists, codec, seg=seg, frac=frac), qs)
            cum, res = ctx.intersect(dix, qs)
            assert np.array_equal(cum, ocum) and np.array_equal(res, ores)
            out = ctx.decompress(dix, list(range(len(DF))))
            assert np.array_equal(out, docs)
        finally:
            dix.close()
        many, off2 = ctx.synth_index_device(51, U, DF, seg=seg, batch_postings=1 << 20, codec=codec, frac=frac)
        try:
            assert np.array_equal(off, off2)
            b = ctx.export(many)
            for name in names:
                assert getattr(b, name).tobytes() == getattr(host, name).tobytes(), name
        finally:
            many.close()
    finally:
        ctx.device_free(d)


@pytest.mark.parametrize("shift", [0, 13, 9])
def test_device_hs_encoder_matches_host_builder(ctx, shift):
    """The HS_SegLRC encoder on the device: the hash tables
    ([shift][nb][start_0 .. start_nb], adaptive or a fixed shift), per-bucket
    fits, bases, offsets and packed residuals byte-identical to the host
    builder's, the index answering queries as the oracle does."""
    d, off, docs = _synth(ctx, 71)
    try:
        host = ea.build_index(docs, off, ea.HS_SEGLRC, seg=128, hs_shift=shift)
        dix = ctx.build_index_device(d, off, codec=ea.HS_SEGLRC, hs_shift=shift)
        try:
            dev = ctx.export(dix)
            for name in ("ind1", "ind2", "seg_num", "seg_offset", "seg_info", "seg_base", "hash_offset", "hash_info"):
                assert getattr(dev, name).tobytes() == getattr(host, name).tobytes(), name
            lists = [docs[int(off[t]):int(off[t + 1])] for t in range(len(DF))]
            rng = np.random.default_rng(6)
            qs = [[int(x) for x in rng.choice(len(DF), int(rng.integers(1, 5)), replace=False)] for _ in range(200)]
            ocum, ores = oracle.eval_queries(oracle.build_index(lists, oracle.HS, seg=128, hs_shift=shift), qs)
            cum, res = ctx.intersect(dix, qs)
            assert np.array_equal(cum, ocum) and np.array_equal(res, ores)
            assert np.array_equal(ctx.decompress(dix, list(range(len(DF)))), docs)
        finally:
            dix.close()
        many, off2 = ctx.synth_index_device(71, U, DF, batch_postings=1 << 20, codec=ea.HS_SEGLRC, hs_shift=shift)
        try:
            b = ctx.export(many)
            for name in ("ind2", "seg_offset", "seg_info", "hash_info"):
                assert getattr(b, name).tobytes() == getattr(host, name).tobytes(), name
        finally:
            many.close()
    finally:
        ctx.device_free(d)


def test_device_encoder_refuses_other_codecs(ctx):
    d, off, docs = _synth(ctx, 61, df=np.array([10, 100], np.uint32))
    try:
        for codec in (ea.NEWPFD, ea.RAW, ea.RAW_HS):
            with pytest.raises(ea.ExcError) as e:
                ctx.build_index_device(d, off, seg=128, codec=codec)
            assert e.value.code == -1
    finally:
        ctx.device_free(d)
