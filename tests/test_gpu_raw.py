"""GPU tests of the uncompressed path (EXC_RAW, the reference's
src/gpu/Uncompressed programs): sorted docID lists stored as they are,
probed by binary search (mqSearch, GU/cuda_binarySearch) or interpolation
search (GU/cuda_interpolationSearch); every result equals the oracle's
exact intersection, from host arrays and from .ind1/.ind2 files."""
import numpy as np
import pytest

import exclucene_amd as ea
import oracle

pytestmark = pytest.mark.gpu


def _lists(seed):
    rng = np.random.default_rng(seed)
    lens = [1, 2, 3, 511, 512, 513, 1000, 5000, 40000, 300000, 2000000]
    lists = [np.sort(rng.choice(1 << 24, n, replace=False)).astype(np.uint32) for n in lens]
    lists.append(np.arange(0, 1 << 21, dtype=np.uint32))            # dense run: r - t == right - left
    lists.append(np.array([0, 5, 6, 7, (1 << 31) - 1], np.uint32))   # extremes
    g = rng.integers(1, 4, 100000).astype(np.uint64)                  # clustered with far jumps:
    g[::997] = rng.integers(1 << 14, 1 << 18, len(g[::997]))          # interpolation guesses far off
    lists.append(np.cumsum(g).astype(np.uint32))
    return lists


@pytest.fixture(scope="module")
def raw(ctx):
    lists = _lists(3)
    docs, off = ea.lists_to_flat(lists)
    built = ea.build_index(docs, off, ea.RAW)
    dix = ctx.upload(built)
    rng = np.random.default_rng(8)
    qs = [[int(x) for x in rng.choice(len(lists), int(rng.integers(1, 6)), replace=False)] for _ in range(400)]
    qs += [[i, i] for i in range(len(lists))] + [[0, len(lists) - 1], [len(lists) - 2, len(lists) - 3]]
    ocum, ores = oracle.eval_queries(oracle.build_index(lists, oracle.SEGLRC, seg=128), qs)
    yield lists, docs, off, built, dix, qs, ocum, ores
    dix.close()
    built.close()


@pytest.mark.parametrize("search", [0, 1, 3], ids=["binary", "regression", "interpolation"])
@pytest.mark.parametrize("persistent", [0, 1])
def test_raw_queries_exact(ctx, raw, search, persistent):
    lists, docs, off, built, dix, qs, ocum, ores = raw
    print("case raw search=%d persistent=%d" % (search, persistent), flush=True)
    ctx.set_option(ea.OPT_SEARCH, search)
    ctx.set_option(ea.OPT_PERSISTENT, persistent)
    try:
        cum, res = ctx.intersect(dix, qs)
    finally:
        ctx.set_option(ea.OPT_SEARCH, 1)
        ctx.set_option(ea.OPT_PERSISTENT, 0)
    assert np.array_equal(cum, ocum) and np.array_equal(res, ores)


def test_raw_files_decompress_and_autotune(ctx, raw, tmp_path):
    lists, docs, off, built, dix, qs, ocum, ores = raw
    assert np.array_equal(ctx.decompress(dix, np.arange(len(lists))), docs[:int(off[-1])])
    built.save(str(tmp_path), "raw")
    lix = ctx.load_files(ea.RAW, 0, str(tmp_path), "raw")
    try:
        t = ctx.autotune(lix, qs[:200], reps=1)
        assert all(s in (0, 1, 3) for s in t["search"])
        cum, res = ctx.intersect(lix, qs)
        assert np.array_equal(cum, ocum) and np.array_equal(res, ores)
    finally:
        ctx.set_tuning(None)
        lix.close()


@pytest.mark.parametrize("kind", [ea.MODEL_PLA, ea.MODEL_MLP, ea.MODEL_PLA_Q], ids=["pla", "mlp", "pla_q"])
def test_raw_learned_window_exact(ctx, raw, kind):
    """Learned windows on uncompressed lists: the lists longer than 1,000
    docIDs get a model; a wave searches only the window its model predicts
    for its candidates' span (+- the recorded error), when the list's
    elements around the window show it holds the span, else the whole list.
    Every result equals the oracle's, with the regression window too on the
    same index."""
    lists, docs, off, built, dix, qs, ocum, ores = raw
    st = ctx.train_models(dix, kind=kind, min_len=1000, epochs=60, samples=2048)
    assert st["nmodels"] >= 5
    try:
        for search in (ea.SEARCH_LEARNED, ea.SEARCH_FIT):
            print("case raw learned kind=%d search=%d" % (kind, search), flush=True)
            ctx.set_option(ea.OPT_SEARCH, search)
            cum, res = ctx.intersect(dix, qs)
            assert np.array_equal(cum, ocum) and np.array_equal(res, ores), search
    finally:
        ctx.set_option(ea.OPT_SEARCH, 1)


def test_raw_unsorted_list_refused(ctx):
    docs = np.array([1, 5, 3, 9], np.uint32)
    off = np.array([0, 4], np.uint64)
    with pytest.raises(ea.ExcError):
        ea.build_index(docs, off, ea.RAW)  # the builder checks order first
    b = ea.build_index(np.array([1, 3, 5, 9], np.uint32), off, ea.RAW)
    try:
        b.ind2[2] = 2  # now 1, 3, 2, 9
        with pytest.raises(ea.ExcError) as e:
            ctx.upload(b)
        assert e.value.code == -4
    finally:
        b.close()


# ---------------------------------------------------------------- hash segments (EXC_RAW_HS)

def _hs_lists(seed):
    lists = _lists(seed)
    rng = np.random.default_rng(seed + 100)
    # one-bucket tables whose maximum is an exact power of two: the maximum's
    # bucket (max >> bits = 1) lies past the table, where the reference reads
    # the next list's table
    lists.append(np.array([3, 1 << 29], np.uint32))
    lists.append(np.sort(np.append(rng.choice(1 << 24, 99, replace=False), 1 << 24)).astype(np.uint32))
    # a long list whose docIDs stop early: most probes lie above its last bucket
    lists.append(np.sort(rng.choice(1 << 16, 3000, replace=False)).astype(np.uint32))
    lists.append(np.zeros(0, np.uint32))
    return lists


@pytest.fixture(scope="module", params=[0, 13], ids=["adaptive", "shift13"])
def raw_hs(ctx, request):
    lists = _hs_lists(4)
    docs, off = ea.lists_to_flat(lists)
    built = ea.build_index(docs, off, ea.RAW_HS, hs_shift=request.param)
    dix = ctx.upload(built)
    rng = np.random.default_rng(9)
    nl = len(lists)
    qs = [[int(x) for x in rng.choice(nl, int(rng.integers(1, 6)), replace=False)] for _ in range(400)]
    # lists nl-4 .. nl-1: [3, 2^29], 99 docIDs + 2^24, the early-stopping list, the empty list
    qs += [[i, i] for i in range(nl)] + [[nl - 4, j] for j in range(nl)] + [[j, nl - 3] for j in range(nl)]
    qs += [[nl - 2, j] for j in range(nl)] + [[j, nl - 2] for j in range(nl)] + [[nl - 1, 0], [0, nl - 1]]
    ocum, ores = oracle.eval_queries(oracle.build_index(lists, oracle.SEGLRC, seg=128), qs)
    yield lists, docs, off, built, dix, qs, ocum, ores, request.param
    dix.close()
    built.close()


@pytest.mark.parametrize("search", [0, 1])
@pytest.mark.parametrize("persistent", [0, 1])
def test_raw_hs_queries_exact(ctx, raw_hs, search, persistent):
    """Every query through the bucket probe equals the oracle's exact
    intersection: binary (the reference's) and interpolation search inside
    the bucket, both launch modes, including docIDs that fall past a list's
    last bucket and the maxima of one-bucket lists."""
    lists, docs, off, built, dix, qs, ocum, ores, shift = raw_hs
    print("case raw_hs shift=%d search=%d persistent=%d" % (shift, search, persistent), flush=True)
    ctx.set_option(ea.OPT_SEARCH, search)
    ctx.set_option(ea.OPT_PERSISTENT, persistent)
    try:
        cum, res = ctx.intersect(dix, qs)
    finally:
        ctx.set_option(ea.OPT_SEARCH, 1)
        ctx.set_option(ea.OPT_PERSISTENT, 1)
    assert np.array_equal(cum, ocum) and np.array_equal(res, ores)
    # the maximum of the [3, 2^29] list is found although its bucket is past its table
    nl = len(lists)
    c, r = ctx.intersect(dix, [[nl - 4, nl - 4], [nl - 3, nl - 3]])
    assert list(r[:c[0]]) == [3, 1 << 29]
    assert np.array_equal(r[c[0]:], lists[nl - 3]) and r[-1] == 1 << 24


def test_raw_hs_files_decompress_and_autotune(ctx, raw_hs, tmp_path):
    lists, docs, off, built, dix, qs, ocum, ores, shift = raw_hs
    assert np.array_equal(ctx.decompress(dix, np.arange(len(lists))), docs[:int(off[-1])])
    built.save(str(tmp_path), "at")
    lix = ctx.load_files(ea.RAW_HS, 0, str(tmp_path), "at")
    try:
        t = ctx.autotune(lix, qs[:200], reps=1)
        assert all(s in (0, 1, 3) for s in t["search"])
        cum, res = ctx.intersect(lix, qs)
        assert np.array_equal(cum, ocum) and np.array_equal(res, ores)
    finally:
        ctx.set_tuning(None)
        lix.close()


def test_raw_hs_malformed_tables_refused(ctx):
    """Tables that would send a probe outside its list are refused at upload:
    starts that decrease, starts that never reach the list length, a bucket
    that does not bracket its docIDs."""
    L = np.sort(np.random.default_rng(5).choice(1 << 20, 2000, replace=False)).astype(np.uint32)
    docs, off = ea.lists_to_flat([L])
    for what in ("decrease", "short", "bracket"):
        b = ea.build_index(docs, off, ea.RAW_HS)
        try:
            h = b.hash_info
            assert len(h) > 6
            if what == "decrease":  # h = [shift, start_0, start_1, start_2, ...]
                h[3] = h[4] + 1
            elif what == "short":
                h[len(h) - 1] = len(L) - 1
            else:
                h[3] += 1
            with pytest.raises(ea.ExcError) as e:
                ctx.upload(b)
            assert e.value.code == -4, what
        finally:
            b.close()
