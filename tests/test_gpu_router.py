"""GPU tests of the router (exc_router_*): one batch routed by ratio class
over several layouts of the same lists (segment sizes, hash-bucket widths,
codecs), each candidate's sub-batch on a stream of its own.  Every route table
gives the oracle's results in the caller's query order, eagerly and replayed
from one captured graph; the autotuner's routes are never slower, on its own
measurements, than the best candidate alone."""
import numpy as np
import pytest

import exclucene_amd as ea
import oracle
from test_gpu_autotune import _classes, _corpus, _ratio_queries

pytestmark = pytest.mark.gpu

# (codec, seg, frac, hs_shift): two segment sizes, the adaptive and a fixed
# bucket width, a gap codec and the one-fit codec
CANDS = [(ea.SEGLRC, 128, 0.0, 0), (ea.SEGLRC, 256, 0.0, 0), (ea.HS_SEGLRC, 128, 0.0, 0),
         (ea.HS_SEGLRC, 128, 0.0, 15), (ea.PARAPFD, 128, 0.1, 0), (ea.LRC, 256, 0.0, 0)]


@pytest.fixture(scope="module")
def routed(ctx):
    lists = _corpus()
    docs, off = ea.lists_to_flat(lists)
    built = [ea.build_index(docs, off, c, seg=s, frac=f, hs_shift=h) for c, s, f, h in CANDS]
    dixs = [ctx.upload(b) for b in built]
    qs = _ratio_queries(lists, 600, 5)
    qs += [[3], [7, 7], [0, len(lists) - 1], [len(lists) - 2, 1, 4], []]
    assert len(set(_classes(lists, qs))) == 6
    ocum, ores = oracle.eval_queries(oracle.build_index(lists, oracle.SEGLRC, seg=128), qs)
    router = ea.Router(ctx, dixs)
    yield lists, qs, ocum, ores, router, dixs
    router.close()
    for d in dixs:
        d.close()
    for b in built:
        b.close()


@pytest.mark.parametrize("graph", [0, 1])
def test_any_route_table_is_exact(routed, graph):
    lists, qs, ocum, ores, router, dixs = routed
    router.set_option(ea.OPT_GRAPH, graph)
    try:
        for trial in range(5):
            rng = np.random.default_rng(trial)
            routes = [int(x) for x in rng.integers(0, len(dixs), ea.RATIO_CLASSES)]
            print("case routes=%s graph=%d" % (routes, graph), flush=True)
            router.set_routes(routes)
            router.set_option(ea.OPT_PERSISTENT, trial % 2)
            cum, res = router.intersect(qs)
            assert np.array_equal(cum, ocum) and np.array_equal(res, ores), routes
    finally:
        router.set_option(ea.OPT_GRAPH, 0)


def test_graph_replays_match(ctx, routed):
    """One plan, the captured routed run replayed with eager work and
    synchronisations between replays; every replay's results are checked."""
    lists, qs, ocum, ores, router, dixs = routed
    router.set_routes([0, 2, 3, 4, 5, 1])
    router.set_option(ea.OPT_GRAPH, 1)
    try:
        router.plan(qs)
        for rep in range(4):
            print("case graph replay %d" % rep, flush=True)
            ms = router.invoke()
            assert ms > 0
            cum, res = router.fetch()
            assert np.array_equal(cum, ocum) and np.array_equal(res, ores), rep
            c2, r2 = ctx.intersect(dixs[rep % len(dixs)], qs[:50])  # eager work in between
            assert np.array_equal(c2, ocum[:50])
    finally:
        router.set_option(ea.OPT_GRAPH, 0)


def test_autotune_routes(routed):
    lists, qs, ocum, ores, router, dixs = routed
    sample = qs[:400]
    t = router.autotune(sample, reps=2)
    print("routes", t, flush=True)
    n = len(dixs)
    assert len(t["single_ms"]) == n and all(x > 0 for x in t["single_ms"])
    assert all(0 <= r < n for r in t["route"])
    assert t["routed_ms"] > 0
    # either the per-class routes won the measured run, or every class went
    # to the fastest candidate alone
    best = int(np.argmin(t["single_ms"]))
    assert t["routed_ms"] < min(t["single_ms"]) or t["route"] == [best] * ea.RATIO_CLASSES
    for graph in (0, 1):
        router.set_option(ea.OPT_GRAPH, graph)
        cum, res = router.intersect(qs)
        assert np.array_equal(cum, ocum) and np.array_equal(res, ores), graph
    router.set_option(ea.OPT_GRAPH, 0)


def test_router_errors(ctx, routed):
    lists, qs, ocum, ores, router, dixs = routed
    with pytest.raises(ea.ExcError):
        router.set_routes([0, 0, 0, 0, 0, len(dixs)])
    other = ea.build_index(*ea.lists_to_flat(lists[:-1]), ea.SEGLRC, seg=128)
    odix = ctx.upload(other)
    try:
        with pytest.raises(ea.ExcError, match="other lists"):
            ea.Router(ctx, [dixs[0], odix])
    finally:
        odix.close()
        other.close()
    fresh = ea.Router(ctx, dixs[:2])
    try:
        with pytest.raises(ea.ExcError):
            fresh.invoke()  # nothing planned
        with pytest.raises(ea.ExcError):
            fresh.plan([[0, len(lists)]])  # term out of range
        fresh.plan([])
        assert fresh.invoke() == 0
        cum, res = fresh.fetch()
        assert len(cum) == 0 and len(res) == 0
    finally:
        fresh.close()
