"""GPU tests of the continuous-batching server: every query served in a mix of
overlapping batches returns exactly what it returns alone, under a tight
device-memory budget too, and errors reach the caller."""
import threading

import numpy as np
import pytest

import exclucene_amd as ea
import oracle

pytestmark = pytest.mark.gpu


def _corpus(seed=41, nl=60):
    rng = np.random.default_rng(seed)
    lens = np.unique(np.concatenate([[1, 7, 200], (10 ** rng.uniform(1.5, 5.7, nl)).astype(int)]))
    return [np.sort(rng.choice(1 << 24, int(n), replace=False)).astype(np.uint32) for n in lens]


def _queries(nl, nq, seed):
    rng = np.random.default_rng(seed)
    return [[int(x) for x in rng.choice(nl, int(rng.integers(1, 5)), replace=False)] for _ in range(nq)]


@pytest.fixture(scope="module")
def served(ctx):
    lists = _corpus()
    docs, off = ea.lists_to_flat(lists)
    built = ea.build_index(docs, off, ea.SEGLRC, seg=128)
    dix = ctx.upload(built)
    yield lists, built, dix
    dix.close()


@pytest.mark.parametrize("nslots,budget", [(1, 0), (3, 0), (4, 1 << 20)])
def test_alone_vs_in_mix(ctx, served, nslots, budget):
    lists, built, dix = served
    qs = _queries(len(lists), 600, 7 + nslots)
    alone = [ctx.intersect(dix, [q])[1] for q in qs]
    ocum, ores = oracle.eval_queries(oracle.build_index(lists, oracle.SEGLRC, seg=128), qs)
    assert np.array_equal(np.concatenate(alone), ores)
    # Poisson arrivals at 20k queries/s
    rng = np.random.default_rng(3)
    arrival = np.cumsum(rng.exponential(1 / 20000.0, len(qs)))
    cap = sum(len(a) for a in alone)
    srv = ea.Server(ctx, dix, nslots=nslots, max_batch=128, mem_budget=budget)
    try:
        lat, cnt, res = srv.replay(qs, arrival, cap)
        st = srv.stats()
    finally:
        srv.close()
    assert np.array_equal(cnt, [len(a) for a in alone])
    assert np.array_equal(res, np.concatenate(alone))
    assert np.all(lat > 0) and st["queries"] == len(qs)
    assert st["batches"] >= (len(qs) + 127) // 128


def test_batches_overlap_on_the_device(ctx, served):
    """Queries that are all waiting when the loop starts fill every free
    slot in one pass, so later batches launch while earlier ones are still in
    flight; each query's docIDs are still those it gets alone."""
    lists, built, dix = served
    qs = _queries(len(lists), 400, 21)
    alone = [ctx.intersect(dix, [q])[1] for q in qs]
    srv = ea.Server(ctx, dix, nslots=3, max_batch=16)
    try:
        lat, cnt, res = srv.replay(qs, np.zeros(len(qs)), sum(len(a) for a in alone))
        st = srv.stats()
    finally:
        srv.close()
    assert np.array_equal(res, np.concatenate(alone))
    assert st["max_inflight_batches"] >= 2
    assert st["overlapped_launches"] > 0
    assert st["batches"] >= len(qs) // 16


def test_submit_wait_from_threads(ctx, served):
    lists, built, dix = served
    qs = _queries(len(lists), 200, 99)
    alone = [ctx.intersect(dix, [q])[1] for q in qs]
    srv = ea.Server(ctx, dix, nslots=2, max_batch=32)
    out = [None] * len(qs)
    try:
        def worker(lo, hi):
            tks = [(i, srv.submit(qs[i])) for i in range(lo, hi)]
            for i, tk in tks:
                out[i] = srv.wait(tk)[0]

        th = [threading.Thread(target=worker, args=(k * 50, (k + 1) * 50)) for k in range(4)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        with pytest.raises(ea.ExcError) as e:
            srv.submit([0, len(lists) + 5])
        assert e.value.code == -1
        with pytest.raises(ea.ExcError):
            srv.wait(123456789)
    finally:
        srv.close()
    for i in range(len(qs)):
        assert np.array_equal(out[i], alone[i]), i


def test_counts_only_server(ctx, served):
    lists, built, dix = served
    qs = _queries(len(lists), 300, 77)
    alone = [len(ctx.intersect(dix, [q])[1]) for q in qs]
    srv = ea.Server(ctx, dix, nslots=3, max_batch=64, counts_only=True)
    try:
        lat, cnt, res = srv.replay(qs, np.zeros(len(qs)), 0)
        tk = srv.submit(qs[0])
        got, _, n = srv.wait(tk, cap=0)
    finally:
        srv.close()
    assert np.array_equal(cnt, alone) and len(res) == 0 and len(got) == 0 and n == alone[0]


@pytest.mark.parametrize("codec,seg,frac", [(ea.LRC, 256, 0.0), (ea.HS_SEGLRC, 128, 0.0), (ea.PARAPFD, 128, 0.1),
                                            (ea.RAW, 128, 0.0)])
def test_server_other_codecs(ctx, codec, seg, frac):
    """The serving loop over every probe codec: each query's docIDs in the mix
    equal its docIDs alone."""
    lists = _corpus(43, 40)
    docs, off = ea.lists_to_flat(lists)
    built = ea.build_index(docs, off, codec, seg=seg, frac=frac)
    dix = ctx.upload(built)
    try:
        qs = _queries(len(lists), 200, 5)
        alone = [ctx.intersect(dix, [q])[1] for q in qs]
        srv = ea.Server(ctx, dix, nslots=3, max_batch=32)
        try:
            lat, cnt, res = srv.replay(qs, np.cumsum(np.full(len(qs), 1e-5)), sum(len(a) for a in alone))
        finally:
            srv.close()
        assert np.array_equal(cnt, [len(a) for a in alone])
        assert np.array_equal(res, np.concatenate(alone))
    finally:
        dix.close()
        built.close()
