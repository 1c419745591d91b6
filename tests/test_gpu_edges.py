"""GPU edge cases across every codec: empty lists inside an index, docIDs up
to 2^31 - 1 (the largest the builders accept), lists of one docID, 0 and the
maximum in one list, and queries mixing them (repeated terms, empty results),
against the oracle; whole lists decode back for all six codecs."""
import itertools

import numpy as np
import pytest

import exclucene_amd as ea
import oracle

pytestmark = pytest.mark.gpu

TOP = (1 << 31) - 1


def _lists():
    rng = np.random.default_rng(5)
    return [np.zeros(0, np.uint32), np.array([TOP], np.uint32),
            np.sort(rng.choice(1 << 31, 3000, replace=False)).astype(np.uint32),
            np.arange(TOP - 2000, TOP + 1, dtype=np.uint64).astype(np.uint32),
            np.array([0, TOP], np.uint32),
            np.sort(rng.choice(1 << 30, 500, replace=False) + (1 << 30)).astype(np.uint32),
            np.zeros(0, np.uint32),
            np.sort(np.concatenate([np.arange(TOP - 100, TOP + 1, dtype=np.uint64).astype(np.uint32),
                                    rng.choice(1 << 20, 4000, replace=False).astype(np.uint32)]))]


CODECS = [(ea.LRC, 256, 0.0), (ea.SEGLRC, 128, 0.0), (ea.SEGLRC, 256, 0.0), (ea.HS_SEGLRC, 128, 0.0),
          (ea.PARAPFD, 128, 0.1), (ea.NEWPFD, 256, 0.1), (ea.RAW, 128, 0.0)]


@pytest.mark.parametrize("codec,seg,frac", CODECS)
def test_edges_match_oracle(ctx, codec, seg, frac):
    lists = _lists()
    docs, off = ea.lists_to_flat(lists)
    b = ea.build_index(docs, off, codec, seg=seg, frac=frac)
    print("case edges codec=%d seg=%d" % (codec, seg), flush=True)
    dix = ctx.upload(b)
    try:
        out = ctx.decompress(dix, np.arange(len(lists)))
        assert np.array_equal(out, np.concatenate(lists))
        if codec == ea.NEWPFD:
            return
        nl = len(lists)
        qs = [[i] for i in range(nl)] + [list(p) for p in itertools.permutations(range(nl), 2)]
        qs += [list(t) for t in itertools.combinations(range(nl), 3)] + [[2, 2, 7], [7, 3, 1], [0], [6, 0]]
        ocum, ores = oracle.eval_queries(oracle.build_index(lists, oracle.SEGLRC, seg=128), qs)
        for search in (0, 1):
            ctx.set_option(ea.OPT_SEARCH, search)
            cum, res = ctx.intersect(dix, qs)
            assert np.array_equal(cum, ocum) and np.array_equal(res, ores), search
    finally:
        ctx.set_option(ea.OPT_SEARCH, 1)
        dix.close()
        b.close()
