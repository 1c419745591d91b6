"""GPU parity: the HIP decode + intersect path through the C ABI against the
CPU oracle (bit-exact docID sets, cumulative counts and checksums)."""
import numpy as np
import pytest

import exclucene_amd as ea
import oracle

pytestmark = pytest.mark.gpu

LENGTHS = [1, 2, 7, 64, 127, 128, 129, 255, 256, 257, 1000, 4000, 20000, 150000, 600000, 64000, 16384]


def _lists(seed, universe=1 << 24):
    rng = np.random.default_rng(seed)
    out = [np.sort(rng.choice(universe, n, replace=False)).astype(np.uint32) for n in LENGTHS]
    # a dense run and a list entirely above / below the others
    out.append(np.arange(5000, 9000, dtype=np.uint32))
    out.append(np.arange(universe - 300, universe, dtype=np.uint32))
    out.append(np.arange(0, 50, dtype=np.uint32))
    return out


def _queries(nl, seed):
    rng = np.random.default_rng(seed)
    qs = [[i] for i in range(nl)]
    qs += [[i, j] for i in range(nl) for j in range(nl) if i != j and (i + j) % 3 == 0]
    qs += [[int(x) for x in rng.choice(nl, int(rng.integers(3, 6)), replace=False)] for _ in range(40)]
    qs += [[13, 13], [14, 13, 12, 11], [nl - 1, nl - 2], [nl - 3, 14], [], [0, 14, 14], [10, 15], [5, 16], [4, 16], [6, 15, 14]]
    return qs


CASES = [(ea.LRC, 256, 0.0, 0), (ea.LRC, 128, 0.0, 0), (ea.SEGLRC, 128, 0.0, 0), (ea.SEGLRC, 256, 0.0, 0),
         (ea.HS_SEGLRC, 128, 0.0, 0), (ea.HS_SEGLRC, 128, 0.0, 13), (ea.PARAPFD, 128, 0.1, 0),
         (ea.PARAPFD, 256, 0.1, 0), (ea.PARAPFD, 128, 0.0, 0),
         # more than 64 exceptions per block: patched in several rounds
         (ea.PARAPFD, 256, 0.5, 0), (ea.PARAPFD, 128, 0.6, 0)]


@pytest.mark.parametrize("codec,seg,frac,shift", CASES)
def test_intersect_matches_oracle(ctx, codec, seg, frac, shift):
    lists = _lists(11)
    docs, off = ea.lists_to_flat(lists)
    built = ea.build_index(docs, off, codec, seg=seg, frac=frac, hs_shift=shift)
    ref = oracle.build_index(lists, codec, seg=seg, frac=frac, hs_shift=shift)
    qs = _queries(len(lists), 3)
    ocum, ores = oracle.eval_queries(ref, qs)
    dix = ctx.upload(built)
    try:
        modes = [(1, 0), (64, 0), (0, 0), (0, 2), (0, 16)] if codec in (ea.LRC, ea.SEGLRC) else [(0, 2), (0, 64)]
        for search in (0, 1):
            for tpe, coop in modes:
                for persistent in (0, 1):
                    ctx.set_option(ea.OPT_SEARCH, search)
                    ctx.set_option(ea.OPT_TPE, tpe)
                    ctx.set_option(ea.OPT_COOP_MAX, coop)
                    ctx.set_option(ea.OPT_PERSISTENT, persistent)
                    print("case codec=%d seg=%d search=%d tpe=%d coop=%d persistent=%d"
                          % (codec, seg, search, tpe, coop, persistent), flush=True)
                    cum, res = ctx.intersect(dix, qs)
                    assert np.array_equal(cum, ocum), (search, tpe, persistent)
                    assert np.array_equal(res, ores), (search, tpe, persistent)
    finally:
        ctx.set_option(ea.OPT_TPE, 0)
        ctx.set_option(ea.OPT_COOP_MAX, 64)
        ctx.set_option(ea.OPT_SEARCH, 1)
        ctx.set_option(ea.OPT_PERSISTENT, 1)
        dix.close()


@pytest.mark.parametrize("codec,seg,frac,shift", CASES)
def test_decompress_roundtrip(ctx, codec, seg, frac, shift):
    lists = _lists(12)
    docs, off = ea.lists_to_flat(lists)
    built = ea.build_index(docs, off, codec, seg=seg, frac=frac, hs_shift=shift)
    dix = ctx.upload(built)
    try:
        terms = list(range(len(lists)))[::-1]
        out = ctx.decompress(dix, terms)
        assert np.array_equal(out, np.concatenate([lists[t] for t in terms]))
    finally:
        dix.close()


def test_c1_plain_query_on_device(ctx):
    """C1 on the device: LRC-256, the plain method (binary search, one lane per
    element, one launch) and the tuned batch path give the oracle's answer."""
    from exclucene_amd import workloads as W
    a, b = W.c1_lists(1)
    docs, off = ea.lists_to_flat([a, b])
    built = ea.build_index(docs, off, ea.LRC, seg=256)
    dix = ctx.upload(built)
    want = np.intersect1d(a, b)
    try:
        for search, tpe, pers in ((0, 1, 0), (1, 0, 1), (1, 64, 0)):
            ctx.set_option(ea.OPT_SEARCH, search)
            ctx.set_option(ea.OPT_TPE, tpe)
            ctx.set_option(ea.OPT_PERSISTENT, pers)
            cum, res = ctx.intersect(dix, [[0, 1]])
            assert cum[0] == len(want) and np.array_equal(res, want), (search, tpe, pers)
    finally:
        ctx.set_option(ea.OPT_SEARCH, 1)
        ctx.set_option(ea.OPT_TPE, 0)
        ctx.set_option(ea.OPT_PERSISTENT, 1)
        dix.close()


@pytest.mark.parametrize("frac", [0.1, 0.6])
def test_parapfd_sparse_and_distinct_paths_agree(ctx, frac):
    """High-ratio ParaPFD probes take the half-block path when the index
    carries its block medians (.segMedian) and the distinct-block rounds when
    it does not (an index written without them): both equal the oracle, in
    both launch modes, on queries of one short list against long ones."""
    lists = _lists(7)
    docs, off = ea.lists_to_flat(lists)
    built = ea.build_index(docs, off, ea.PARAPFD, seg=128, frac=frac)
    ref = oracle.build_index(lists, ea.PARAPFD, seg=128, frac=frac)
    n = [len(L) for L in lists]
    shorts = [i for i in range(len(lists)) if 2 <= n[i] <= 1000]
    longs = [i for i in range(len(lists)) if n[i] >= 20000]
    qs = [[s, t] for s in shorts for t in longs] + [[s, t, u] for s in shorts for t in longs for u in longs if t < u]
    ocum, ores = oracle.eval_queries(ref, qs)
    with_medians = ctx.upload(built)
    view = ea.IndexHost.from_buffer_copy(built.view)
    view.seg_median = None
    no_medians = ea.DeviceIndex(ctx, _upload_view(ctx, view), lengths=np.diff(off).astype(np.uint32))
    try:
        for dix in (with_medians, no_medians):
            for persistent in (0, 1):
                ctx.set_option(ea.OPT_PERSISTENT, persistent)
                cum, res = ctx.intersect(dix, qs)
                assert np.array_equal(cum, ocum) and np.array_equal(res, ores), (dix is no_medians, persistent)
    finally:
        ctx.set_option(ea.OPT_PERSISTENT, 1)
        with_medians.close()
        no_medians.close()
        built.close()


def _upload_view(ctx, view):
    import ctypes
    h = ctypes.c_void_p()
    ea._check(ea.lib().exc_index_upload(ctx.handle, ctypes.byref(view), ctypes.byref(h)))
    return h
