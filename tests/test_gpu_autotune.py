"""GPU tests of per-ratio-class search knobs and the autotuner: any table of
knobs gives the oracle's results, and the tuner's choice is never slower than
the untuned options on its own measurements."""
import numpy as np
import pytest

import exclucene_amd as ea
import oracle

pytestmark = pytest.mark.gpu


def _corpus(seed=51):
    rng = np.random.default_rng(seed)
    lens = [3, 40, 300, 1000, 3000, 10000, 40000, 120000, 400000, 1200000, 3000000]
    lens += list((10 ** rng.uniform(1, 6, 12)).astype(int))
    return [np.sort(rng.choice(1 << 25, int(n), replace=False)).astype(np.uint32) for n in lens]


def _ratio_queries(lists, nq, seed):
    rng = np.random.default_rng(seed)
    n = np.array([len(L) for L in lists])
    out = []
    for _ in range(nq):
        k = int(rng.integers(1, 5))
        out.append([int(x) for x in rng.choice(len(lists), k, replace=False)])
    return out


def _classes(lists, qs):
    n = [len(L) for L in lists]
    cls = []
    for q in qs:
        s = sorted(q, key=lambda t: n[t])
        if len(s) < 2:
            cls.append(5)
            continue
        r, k = n[s[1]] / max(1, n[s[0]]), 0
        while k < 4 and r >= 10:
            r /= 10
            k += 1
        cls.append(k)
    return np.array(cls)


@pytest.mark.parametrize("codec,frac", [(ea.SEGLRC, 0.0), (ea.PARAPFD, 0.1), (ea.HS_SEGLRC, 0.0), (ea.LRC, 0.0)])
def test_any_knob_table_is_exact(ctx, codec, frac):
    lists = _corpus()
    docs, off = ea.lists_to_flat(lists)
    seg = 256 if codec == ea.LRC else 128
    built = ea.build_index(docs, off, codec, seg=seg, frac=frac)
    ref = oracle.build_index(lists, codec, seg=seg, frac=frac)
    qs = _ratio_queries(lists, 500, 3)
    assert len(set(_classes(lists, qs))) >= 5
    ocum, ores = oracle.eval_queries(ref, qs)
    dix = ctx.upload(built)
    coops = [16, 32, 64] if codec == ea.PARAPFD else [0, 2, 8, 64]
    try:
        for trial in range(4):
            rng = np.random.default_rng(trial)
            table = {"coop_max": [int(rng.choice(coops)) for _ in range(6)],
                     "search": [int(rng.integers(0, 2)) for _ in range(6)], "persistent": trial % 2}
            ctx.set_tuning(table)
            cum, res = ctx.intersect(dix, qs)
            assert np.array_equal(cum, ocum) and np.array_equal(res, ores), (codec, table)
    finally:
        ctx.set_tuning(None)
        ctx.set_option(ea.OPT_PERSISTENT, 1)
        dix.close()


def test_autotune_choice(ctx):
    lists = _corpus()
    docs, off = ea.lists_to_flat(lists)
    built = ea.build_index(docs, off, ea.SEGLRC, seg=128)
    ref = oracle.build_index(lists, ea.SEGLRC, seg=128)
    qs = _ratio_queries(lists, 800, 4)
    ocum, ores = oracle.eval_queries(ref, qs)
    dix = ctx.upload(built)
    try:
        t = ctx.autotune(dix, qs, reps=2)
        cls = _classes(lists, qs)
        for k in range(6):
            assert t["queries"][k] == int((cls == k).sum())
            if 0 < t["queries"][k] and k < 5:
                assert t["ms"][k] <= t["ms_default"][k]
                assert t["coop_max"][k] in (0, 2, 4, 8, 16, 32, 64) and t["search"][k] in (0, 1)
        assert t["persistent"] in (0, 1)
        cum, res = ctx.intersect(dix, qs)
        assert np.array_equal(cum, ocum) and np.array_equal(res, ores)
        # a server started from the tuned context uses the same table
        srv = ea.Server(ctx, dix, nslots=2, max_batch=64)
        try:
            lat, cnt, sres = srv.replay(qs, np.zeros(len(qs)), len(ores))
        finally:
            srv.close()
        assert np.array_equal(sres, ores)
    finally:
        ctx.set_tuning(None)
        ctx.set_option(ea.OPT_PERSISTENT, 1)
        dix.close()
