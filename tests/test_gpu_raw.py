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


@pytest.mark.parametrize("search", [0, 1])
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
        assert all(s in (0, 1) for s in t["search"])
        cum, res = ctx.intersect(lix, qs)
        assert np.array_equal(cum, ocum) and np.array_equal(res, ores)
    finally:
        ctx.set_tuning(None)
        lix.close()


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
