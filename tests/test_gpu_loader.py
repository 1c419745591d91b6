"""GPU tests of index loading: an index whose .ind2 spans more than eight
8 MiB staging chunks (so each of the 4 staging threads reuses both of its
buffers)
reaches the device intact through exc_index_upload and through
exc_index_load_files (mapped reference-format files): whole lists decode back
to the source lists, and queries agree with the oracle."""
import numpy as np
import pytest

import exclucene_amd as ea
import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def big():
    rng = np.random.default_rng(12)
    df = np.concatenate([[1, 2, 129], rng.integers(1000, 4_000_000, 34)]).astype(np.uint32)
    docs, off = ea.synth_lists(df, 1 << 26, 12)
    built = ea.build_index(docs, off, ea.SEGLRC, seg=128)
    assert built.ind2.nbytes > 8 * (8 << 20), built.ind2.nbytes
    return df, docs, off, built


def _check(ctx, dix, df, docs, off):
    out = ctx.decompress(dix, np.arange(len(df)))
    assert np.array_equal(out, docs[:int(off[-1])])


def test_staged_upload_and_file_load(ctx, big, tmp_path):
    df, docs, off, built = big
    print("case staged upload", flush=True)
    dix = ctx.upload(built)
    try:
        _check(ctx, dix, df, docs, off)
    finally:
        dix.close()
    built.save(str(tmp_path), "big")
    print("case load_files", flush=True)
    lix = ctx.load_files(ea.SEGLRC, 128, str(tmp_path), "big")
    try:
        _check(ctx, lix, df, docs, off)
        lists = [docs[int(off[t]):int(off[t + 1])] for t in range(len(df))]
        rng = np.random.default_rng(5)
        qs = [[int(x) for x in rng.choice(len(df), int(rng.integers(1, 5)), replace=False)] for _ in range(60)]
        ocum, ores = oracle.eval_queries(oracle.build_index(lists, oracle.SEGLRC, seg=128), qs)
        cum, res = ctx.intersect(lix, qs)
        assert np.array_equal(cum, ocum) and np.array_equal(res, ores)
    finally:
        lix.close()


def test_load_files_errors(ctx, tmp_path):
    with pytest.raises(ea.ExcError) as e:
        ctx.load_files(ea.SEGLRC, 128, str(tmp_path), "missing")
    assert e.value.code == -5  # EXC_ERR_IO
