"""GPU tests of the drop-in surface: the reference's stage sequence over a
.query byte stream with threshold batching, hipGraph replays, the persistent
work queue, index files written and read back in the reference layout, and
error reporting instead of exit()."""
import numpy as np
import pytest

import exclucene_amd as ea
import oracle

pytestmark = pytest.mark.gpu


def _corpus(seed=21, nl=40):
    rng = np.random.default_rng(seed)
    lens = np.unique(np.concatenate([[1, 5, 300], (10 ** rng.uniform(1, 5.5, nl)).astype(int)]))
    return [np.sort(rng.choice(1 << 24, int(n), replace=False)).astype(np.uint32) for n in lens]


def _queries(nl, nq, seed):
    rng = np.random.default_rng(seed)
    return [[int(x) for x in rng.choice(nl, int(rng.integers(1, 6)), replace=False)] for _ in range(nq)]


@pytest.mark.parametrize("codec,frac", [(ea.SEGLRC, 0.0), (ea.PARAPFD, 0.1), (ea.HS_SEGLRC, 0.0), (ea.LRC, 0.0),
                                        (ea.RAW, 0.0)])
def test_stage_sequence_with_thresholds(ctx, codec, frac):
    """Run()-style loop: CPUPreprocess -> htod -> kernelInvoke -> dtoh -> verify
    over the whole stream at several thresholds; every threshold must give the
    same checksums, equal to the oracle's (LRC_Intersection.cu:449-540)."""
    lists = _corpus()
    docs, off = ea.lists_to_flat(lists)
    built = ea.build_index(docs, off, codec, seg=128, frac=frac)
    # the uncompressed lists' oracle: the same intersections over a SegLRC encoding
    ref = oracle.build_index(lists, oracle.SEGLRC if codec == ea.RAW else codec, seg=128, frac=frac)
    qs = _queries(len(lists), 300, 5)
    ocum, ores = oracle.eval_queries(ref, qs)
    ocs = oracle.checksum(ocum, ores)
    stream = ea.queries_to_stream(qs)
    dix = ctx.upload(built)
    try:
        for threshold in (0, 1000, 50_000, 1 << 30):
            pos, cs, batches, allres = 0, [0, 0, 0], 0, []
            while pos < len(stream):
                bi, pos = ctx.CPUPreprocess(dix, stream, pos, threshold)
                ctx.htodTransfer(bi)
                ctx.kernelInvoke(bi)
                cum, res = ctx.dtohTransfer(bi)
                c = ctx.verify(bi, cum, res)
                cs = [cs[0] + c[0], cs[1] + c[1], cs[2] ^ c[2]]
                allres.append(res)
                batches += 1
            assert tuple(cs) == ocs, threshold
            assert np.array_equal(np.concatenate(allres), ores)
            if threshold == 0:
                assert batches == len(qs)  # every query crosses a zero threshold on its own
    finally:
        dix.close()


def test_graph_replays_match_eager(ctx):
    lists = _corpus(22)
    docs, off = ea.lists_to_flat(lists)
    built = ea.build_index(docs, off, ea.SEGLRC, seg=128)
    dix = ctx.upload(built)
    qs = _queries(len(lists), 200, 6)
    try:
        ctx.set_option(ea.OPT_GRAPH, 0)
        ecum, eres = ctx.intersect(dix, qs)
        for persistent in (1, 0):
            ctx.set_option(ea.OPT_PERSISTENT, persistent)
            ctx.set_option(ea.OPT_GRAPH, 1)
            bi = ctx.plan(dix, qs)
            for rep in range(4):
                ctx.kernelInvoke(bi)  # first call captures, later calls replay the same graph
                cum, res = ctx.dtohTransfer(bi)
                assert np.array_equal(cum, ecum) and np.array_equal(res, eres), (persistent, rep)
                # the program's own eager work and a synchronisation between replays
                out = ctx.decompress(dix, [0, 3, 7])
                assert np.array_equal(out, np.concatenate([lists[0], lists[3], lists[7]]))
    finally:
        ctx.set_option(ea.OPT_GRAPH, 0)
        ctx.set_option(ea.OPT_PERSISTENT, 1)
        dix.close()


@pytest.mark.parametrize("codec,frac,shift", [(ea.LRC, 0.0, 0), (ea.SEGLRC, 0.0, 0), (ea.HS_SEGLRC, 0.0, 13),
                                              (ea.PARAPFD, 0.1, 0)])
def test_files_roundtrip_on_device(ctx, tmp_path, codec, frac, shift):
    lists = _corpus(23, 12)
    docs, off = ea.lists_to_flat(lists)
    seg = 256 if codec == ea.LRC else 128
    built = ea.build_index(docs, off, codec, seg=seg, frac=frac, hs_shift=shift)
    built.save(str(tmp_path), "idx")
    dix = ctx.load_files(codec, seg, str(tmp_path), "idx")
    try:
        out = ctx.decompress(ea.DeviceIndex(ctx, dix.handle, built), list(range(len(lists))))
        assert np.array_equal(out, np.concatenate(lists))
    finally:
        dix.close()


def test_errors_are_reported_not_fatal(ctx):
    lists = _corpus(24, 5)
    docs, off = ea.lists_to_flat(lists)
    built = ea.build_index(docs, off, ea.SEGLRC, seg=128)
    dix = ctx.upload(built)
    try:
        with pytest.raises(ea.ExcError) as e:
            ctx.intersect(dix, [[0, 99999]])
        assert e.value.code == -1
        with pytest.raises(ea.ExcError):
            ctx.set_option(ea.OPT_TPE, 7)
        # capacity: result buffer too small
        bi = ctx.plan(dix, [[len(lists) - 1]])
        ctx.kernelInvoke(bi)
        with pytest.raises(ea.ExcError) as e:
            ctx.dtohTransfer(bi, cap=1)
        assert e.value.code == -3
        # still usable afterwards
        cum, res = ctx.intersect(dix, [[0, 1]])
        assert len(cum) == 1
    finally:
        dix.close()


def test_budget_admission():
    c = ea.Context(0, budget=1 << 20)
    try:
        lists = _corpus(25, 50)
        docs, off = ea.lists_to_flat(lists)
        built = ea.build_index(docs, off, ea.SEGLRC, seg=128)
        # the corpus is sized to exceed the 1 MiB budget
        assert built.compressed_bytes() > (1 << 20)
        with pytest.raises(ea.ExcError) as e:
            c.upload(built)
        assert e.value.code == -6
    finally:
        c.close()


def test_decompress_into_device_memory(ctx):
    lists = _corpus(26, 15)
    docs, off = ea.lists_to_flat(lists)
    for codec, frac in ((ea.SEGLRC, 0.0), (ea.PARAPFD, 0.1)):
        built = ea.build_index(docs, off, codec, seg=128, frac=frac)
        dix = ctx.upload(built)
        try:
            terms = [3, 0, 7, 7, 14]
            want = np.concatenate([lists[t] for t in terms])
            cap = len(want) + 10
            out = ctx.device_alloc(cap * 4)
            ctx.to_device(out, np.full(cap, 0xFFFFFFFF, np.uint32))
            n = ctx.decompress_device(dix, terms, out, cap)
            got = ctx.from_device(out, cap)
            assert n == len(want) and np.array_equal(got[:n], want)
            assert np.all(got[n:] == 0xFFFFFFFF)  # nothing written past the output
            assert ctx.last_times()[0] > 0
            with pytest.raises(ea.ExcError) as e:
                ctx.decompress_device(dix, terms, out, len(want) - 1)
            assert e.value.code == -3
            ctx.device_free(out)
            with pytest.raises(ea.ExcError):
                ctx.device_free(out)
        finally:
            dix.close()


def test_scan_over_many_blocks(ctx):
    """A batch of more than three scan blocks of units (4,096 units each): the
    block that finishes last turns the block sums into offsets and leaves the
    ticket at zero, so eager launches, graph replays, a one-block batch and a
    batch without units that follow all give the exact intersections
    (np.intersect1d over the lists themselves)."""
    rng = np.random.default_rng(31)
    lists = [np.sort(rng.choice(1 << 23, int(n), replace=False)).astype(np.uint32)
             for n in rng.integers(400_000, 1_500_000, 12)]
    lists.append(np.zeros(0, np.uint32))
    docs, off = ea.lists_to_flat(lists)
    dix = ctx.upload(ea.build_index(docs, off, ea.SEGLRC, seg=128))

    def expect(qs):
        cum, res, c = [], [], 0
        for q in qs:
            r = lists[q[0]]
            for t in q[1:]:
                r = np.intersect1d(r, lists[t], assume_unique=True)
            res.append(r)
            c += len(r)
            cum.append(c)
        return np.array(cum, np.int64), np.concatenate(res).astype(np.uint32)

    qs = [[int(x) for x in rng.choice(12, int(rng.integers(2, 4)), replace=False)] for _ in range(24)]
    units = sum(-(-min(len(lists[t]) for t in q) // 512) for q in qs)
    assert units > 3 * 4096
    batches = [qs, qs[:3], [[12, 0], [5, 12]]]
    try:
        for graph in (0, 1):
            for persistent in (0, 1):
                ctx.set_option(ea.OPT_GRAPH, graph)
                ctx.set_option(ea.OPT_PERSISTENT, persistent)
                for b in batches:
                    ecum, eres = expect(b)
                    bi = ctx.plan(dix, b)
                    for rep in range(3):
                        ctx.kernelInvoke(bi)
                        cum, res = ctx.dtohTransfer(bi)
                        assert np.array_equal(np.asarray(cum, np.int64), ecum), (graph, persistent, len(b), rep)
                        assert np.array_equal(res, eres), (graph, persistent, len(b), rep)
    finally:
        ctx.set_option(ea.OPT_GRAPH, 0)
        ctx.set_option(ea.OPT_PERSISTENT, 1)
        dix.close()


def test_page_locked_result_buffers(ctx):
    """Results copied into page-locked buffers (exc_host_alloc) equal the
    oracle's, through exc_intersect and through the stages, reused across
    batches; a pointer that is not the context's is refused."""
    lists = _corpus(23)
    docs, off = ea.lists_to_flat(lists)
    built = ea.build_index(docs, off, ea.SEGLRC, seg=128)
    ref = oracle.build_index(lists, oracle.SEGLRC, seg=128)
    dix = ctx.upload(built)
    try:
        batches = [_queries(len(lists), 400, seed) for seed in (6, 7)]
        cap = max(ea._result_cap(dix.ucnt, *ea.queries_to_flat(qs)) for qs in batches)
        cum_buf, res_buf = ctx.host_empty(400), ctx.host_empty(cap)
        for seed, qs in zip((6, 7), batches):
            ocum, ores = oracle.eval_queries(ref, qs)
            cum, res = ctx.intersect(dix, qs, out=(cum_buf, res_buf))
            assert np.array_equal(cum, ocum) and np.array_equal(res, ores), seed
            bi = ctx.plan(dix, qs)
            ctx.kernelInvoke(bi)
            cum, res = ctx.dtohTransfer(bi, out=(cum_buf, res_buf))
            assert np.array_equal(cum, ocum) and np.array_equal(res, ores), seed
        with pytest.raises(ea.ExcError):  # inside a buffer, not its start
            ea._check(ea.lib().exc_host_free(ctx.handle, res_buf[8:].ctypes.data))
    finally:
        dix.close()
