"""exc_intersect_noalloc, the entry point that allocates nothing: through the
C ABI its counts and docIDs equal exc_intersect's on every probe codec and in
both launch modes, on two calls in a row (the second with a smaller batch, so
anything a call leaves in the shared scratch area must not leak into the
next), and batches beyond its fixed scratch area are refused before any work
is queued."""
import numpy as np
import pytest

import exclucene_amd as ea

pytestmark = pytest.mark.gpu


def _corpus(seed, lens):
    rng = np.random.default_rng(seed)
    return [np.sort(rng.choice(1 << 24, int(n), replace=False)).astype(np.uint32) for n in lens]


def _queries(rng, nl, nq, tmax=5):
    return [[int(x) for x in rng.choice(nl, int(rng.integers(1, tmax + 1)), replace=False)] for _ in range(nq)]


@pytest.fixture(scope="module")
def nactx():
    c = ea.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("codec,seg,frac", [(ea.SEGLRC, 128, 0.0), (ea.PARAPFD, 128, 0.1), (ea.LRC, 256, 0.0),
                                            (ea.HS_SEGLRC, 128, 0.0), (ea.RAW, 128, 0.0)])
@pytest.mark.parametrize("persistent", [0, 1])
def test_noalloc_equals_intersect_two_calls(nactx, codec, seg, frac, persistent):
    lens = [1, 7, 128, 129, 600, 3000, 20000, 90000, 250000, 400000, 0]
    lists = _corpus(61, lens)
    docs, off = ea.lists_to_flat(lists)
    built = ea.build_index(docs, off, codec, seg=seg, frac=frac)
    dix = nactx.upload(built)
    rng = np.random.default_rng(62)
    try:
        nactx.set_option(ea.OPT_PERSISTENT, persistent)
        big = _queries(rng, len(lists), 300) + [[], [10, 9], [9]]
        small = _queries(rng, 6, 40, 3)  # shorter lists, fewer queries and units
        for qs in (big, small):
            print("noalloc: codec %d persistent %d, %d queries" % (codec, persistent, len(qs)), flush=True)
            want_cum, want_res = nactx.intersect(dix, qs)
            cum, res = nactx.intersect_noalloc(dix, qs)
            assert np.array_equal(cum, want_cum)
            assert np.array_equal(res, want_res)
        # the context's own planned batch is untouched by the calls: replaying
        # its stages still answers it
        bi = nactx.plan(dix, big)
        nactx.intersect_noalloc(dix, small)
        nactx.kernelInvoke(bi)
        cum, res = nactx.dtohTransfer(bi)
        want_cum, want_res = nactx.intersect(dix, big)
        assert np.array_equal(cum, want_cum) and np.array_equal(res, want_res)
    finally:
        nactx.set_option(ea.OPT_PERSISTENT, 1)
        dix.close()
        built.close()


def test_noalloc_refuses_batches_beyond_its_scratch(nactx):
    lists = _corpus(63, [400000, 500000, 3])
    docs, off = ea.lists_to_flat(lists)
    built = ea.build_index(docs, off, ea.SEGLRC, seg=128)
    dix = nactx.upload(built)
    try:
        # 40 x 400,000 shortest-list docIDs need about 128 MB of stage and results
        with pytest.raises(ea.ExcError) as e:
            nactx.intersect_noalloc(dix, [[0, 1]] * 40)
        assert e.value.code == -3 and "scratch" in str(e.value)
        with pytest.raises(ea.ExcError) as e:
            nactx.intersect_noalloc(dix, [[0, 1, 2] * 22])  # 66 terms
        assert e.value.code == -3
        with pytest.raises(ea.ExcError) as e:
            nactx.intersect_noalloc(dix, [[0, 7]])
        assert e.value.code == -1
        # and a batch that fits still runs after the refusals
        want = nactx.intersect(dix, [[0, 1], [2, 0], [1]])
        got = nactx.intersect_noalloc(dix, [[0, 1], [2, 0], [1]])
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    finally:
        dix.close()
        built.close()
