"""CPU-side checks of the product library: the index builder reproduces the
oracle's encoder byte for byte, the synthetic corpus is well formed and
deterministic, the C ABI exports every declared entry point, and the bench's
algorithmic-byte count agrees with the oracle's."""
import ctypes
import os
import re

import numpy as np
import pytest

import exclucene_amd as ea
import oracle
from exclucene_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def _lists(seed):
    rng = np.random.default_rng(seed)
    out = [np.fromfile(os.path.join(GOLD, "InvertedList.u32"), np.uint32)]
    for n in (1, 2, 63, 64, 127, 128, 129, 256, 257, 3000, 90000):
        out.append(np.sort(rng.choice(1 << 24, n, replace=False)).astype(np.uint32))
    out.append(np.arange(7, 7 + 5000, dtype=np.uint32))
    out.append(np.zeros(0, np.uint32))
    return out


CASES = [(ea.LRC, 256, 0.0, 0), (ea.LRC, 128, 0.0, 0), (ea.SEGLRC, 128, 0.0, 0), (ea.SEGLRC, 256, 0.0, 0),
         (ea.HS_SEGLRC, 128, 0.0, 0), (ea.HS_SEGLRC, 128, 0.0, 13), (ea.PARAPFD, 128, 0.1, 0),
         (ea.PARAPFD, 256, 0.1, 0), (ea.PARAPFD, 256, 0.25, 0), (ea.LRC, 256, 0.1, 0)]


@pytest.mark.parametrize("codec,seg,frac,shift", CASES)
def test_builder_bytes_equal_oracle(codec, seg, frac, shift):
    lists = _lists(1)
    docs, off = ea.lists_to_flat(lists)
    b = ea.build_index(docs, off, codec, seg=seg, frac=frac, hs_shift=shift, nthreads=3)
    o = oracle.build_index(lists, codec, seg=seg, frac=frac, hs_shift=shift)
    S = len(b.seg_offset)
    assert np.array_equal(b.ind2, o.words[:o.total_words])
    assert np.array_equal(b.ind1["m_urlcount"], o.ucnt)
    assert np.array_equal(b.ind1["m_off"], o.uwlist * 4)
    assert np.array_equal(b.seg_num, o.segnum)
    assert np.array_equal(b.seg_offset, o.seg_off[:S])
    if codec != ea.HS_SEGLRC:
        assert np.array_equal(b.seg_head, o.seg_head[:S])
    if codec == ea.PARAPFD:
        assert np.array_equal(b.seg_median, o.seg_median[:S])
    if codec == ea.LRC:
        assert b.lrc_info.tobytes() == o.lrc.tobytes()
    if codec in (ea.SEGLRC, ea.HS_SEGLRC):
        assert b.seg_info.tobytes() == o.fit[:S].tobytes()
        assert np.array_equal(b.seg_base, o.base[:S])
    if codec == ea.SEGLRC:
        assert b.list_info.tobytes() == o.list_fit.tobytes()
    if codec == ea.HS_SEGLRC:
        assert np.array_equal(b.hash_info, o.hash_info)
        assert np.array_equal(b.hash_offset, o.hash_off)


def test_builder_thread_count_invariant():
    docs, off = ea.lists_to_flat(_lists(2))
    a = ea.build_index(docs, off, ea.SEGLRC, nthreads=1)
    b = ea.build_index(docs, off, ea.SEGLRC, nthreads=7)
    assert np.array_equal(a.ind2, b.ind2) and a.seg_info.tobytes() == b.seg_info.tobytes()


def test_builder_rejects_bad_lists():
    with pytest.raises(ea.ExcError):
        ea.build_index(np.array([5, 3], np.uint32), np.array([0, 2], np.uint64), ea.SEGLRC)
    with pytest.raises(ea.ExcError):
        ea.build_index(np.array([5, 5], np.uint32), np.array([0, 2], np.uint64), ea.PARAPFD)
    with pytest.raises(ea.ExcError):
        ea.build_index(np.array([1, 1 << 31], np.uint32), np.array([0, 2], np.uint64), ea.LRC)


def test_file_layout_sizes(tmp_path):
    lists = _lists(3)
    docs, off = ea.lists_to_flat(lists)
    b = ea.build_index(docs, off, ea.SEGLRC, seg=128)
    b.save(str(tmp_path), "idx")
    nt = len(lists)
    sz = {p.name.split(".", 1)[1]: p.stat().st_size for p in tmp_path.iterdir()}
    assert sz["ind1"] == 16 * nt                      # {u32 count; u64 offset}, 16 B
    assert sz["ind2"] == 4 * len(b.ind2)
    assert sz["segNum"] == 4 * nt
    assert sz["segOffset"] == sz["segHead"] == 4 * b.nseg
    assert sz["regression_info"] == 16 * b.nseg       # {slope, icpt, rangeL, rangeR}
    assert sz["ind2_distance_base"] == 4 * b.nseg
    assert sz["regression_info_list"] == 16 * nt
    raw = np.fromfile(tmp_path / "idx.ind1", ea.IND1)
    assert np.array_equal(raw["m_urlcount"], [len(x) for x in lists])


def test_synth_lists_well_formed_and_deterministic():
    df = np.array([0, 1, 5, 1000, 100000, 1 << 18, (1 << 22) + 3], np.uint32)
    a, off = ea.synth_lists(df, 1 << 25, 42, nthreads=4)
    b, _ = ea.synth_lists(df, 1 << 25, 42, nthreads=1)
    c, _ = ea.synth_lists(df, 1 << 25, 43, nthreads=4)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    for t in range(len(df)):
        x = a[int(off[t]):int(off[t + 1])].astype(np.int64)
        assert len(x) == df[t]
        if len(x):
            assert np.all(np.diff(x) > 0) and x[0] >= 0 and x[-1] < (1 << 25)
    # uniformity of a dense list: mean of U(0, 2^25) within 1%
    x = a[int(off[6]):int(off[7])].astype(np.float64)
    assert abs(x.mean() / (1 << 24) - 1.0) < 0.01


def test_zipf_table_closed_form():
    df = W.zipf_df()
    assert df[0] == 1 << 23 and df[3] == 1 << 23 and df[4] == (1 << 25) // 5
    assert df[-1] == (1 << 25) // 100000
    assert int(df.sum(dtype=np.int64)) == sum(min(1 << 23, (1 << 25) // k) for k in range(1, 100001))


def test_c_abi_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "exclucene_amd.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    names = set(re.findall(r"\b(exc_[A-Za-z_0-9]+)\s*\(", hdr))
    assert len(names) >= 20
    so = ctypes.CDLL(ea.LIB_PATH)
    for n in sorted(names):
        assert hasattr(so, n), n
    assert names <= set(ea.EXPORTS) | {n for n in names if hasattr(so, n)}


def test_algorithmic_bytes_agree_with_oracle():
    import bench
    lists = _lists(4)[:-1]
    docs, off = ea.lists_to_flat(lists)
    rng = np.random.default_rng(5)
    qs = [[int(x) for x in rng.choice(len(lists), int(rng.integers(2, 5)), replace=False)] for _ in range(30)]
    for codec, seg, frac, shift in CASES[:7]:
        b = ea.build_index(docs, off, codec, seg=seg, frac=frac, hs_shift=shift)
        o = oracle.build_index(lists, codec, seg=seg, frac=frac, hs_shift=shift)
        assert bench.algorithmic_bytes(b, docs, off, qs) == oracle.query_bytes(o, qs)


def test_no_gpu_calls_without_device():
    # library loads and reports a clean error (not a crash) when no device is visible
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU may be present")
    with pytest.raises(ea.ExcError):
        ea.Context(0)


@pytest.mark.parametrize("shift", [0, 13, 9])
def test_raw_hash_segment_tables_equal_oracle(shift, tmp_path):
    """The uncompressed hash-segment index: .ind2 is the docIDs, and every
    list's [shift][start_0 .. start_nb] table equals the oracle's restatement
    of the reference's generator; the files carry the reference's names."""
    lists = _lists(5) + [np.array([3, 1 << 29], np.uint32), np.array([1 << 24], np.uint32)]
    docs, off = ea.lists_to_flat(lists)
    b = ea.build_index(docs, off, ea.RAW_HS, hs_shift=shift, nthreads=3)
    assert np.array_equal(b.ind2, np.concatenate(lists))
    ends = np.append(b.hash_offset[1:], len(b.hash_info))
    for t, L in enumerate(lists):
        assert np.array_equal(b.hash_info[b.hash_offset[t]:ends[t]], oracle.gu_hashseg_table(L, shift=shift)), t
    b.save(str(tmp_path), "at")
    names = sorted(p.name for p in tmp_path.iterdir())
    assert "at_HashSeg_info" in names and "at_HashSeg_offset" in names
    assert np.array_equal(np.fromfile(tmp_path / "at_HashSeg_info", np.uint32), b.hash_info)
    assert np.array_equal(np.fromfile(tmp_path / "at_HashSeg_offset", np.uint32), b.hash_offset)


# ---------------------------------------------------------------- uncompressed regression rows

def _gu_lists(seed):
    rng = np.random.default_rng(seed)
    lists = [np.sort(rng.choice(1 << 26, n, replace=False)).astype(np.uint32) for n in (2, 3, 100, 5000, 200000)]
    g = rng.integers(1, 40, 30000).astype(np.uint64)  # a dense head then a sparse tail: a poor line
    g[15000:] *= 500
    lists.append(np.cumsum(g).astype(np.uint32))
    lists.append(np.array([77], np.uint32))
    lists.append(np.zeros(0, np.uint32))
    return lists


def _gu_ranges(L, slope, icpt):
    """Each element's signed distance from the line in float32, as the
    reference's getRegressionInfo takes it: (float)docID, minus the float
    intercept, over the float slope, minus (float)position."""
    x = (L.astype(np.float32) - np.float32(icpt)) / np.float32(slope)
    return x - np.arange(len(L)).astype(np.float32)


def test_gu_regression_rows(tmp_path):
    """The regression rows of an uncompressed index
    (GU/cuda_runRegressionAlgorithm/cuda_linearRegression.cu:792-925):
    slope and intercept agree with numpy's least-squares line, the ranges
    are exactly floor(furthest distance) + 1 on either side (so every element
    lies strictly inside its window), the one-element and empty lists take
    the reference's special rows, and the rows travel through
    <name>_regression_info."""
    lists = _gu_lists(21)
    docs, off = ea.lists_to_flat(lists)
    b = ea.build_index(docs, off, ea.RAW, nthreads=2)
    try:
        g = b.gu_regression
        assert g is not None and len(g) == len(lists)
        for t, L in enumerate(lists):
            row = g[t]
            if len(L) == 1:
                assert (row["fSlope"], row["fIntercept"], row["nRangeLeft"], row["nRangeRight"]) == (1.0, 77.0, 0, 0)
                continue
            if len(L) == 0:
                assert np.isnan(row["fSlope"]) and row["nRangeLeft"] == 1 and row["nRangeRight"] == 1
                continue
            s, c = np.polyfit(np.arange(len(L), dtype=np.float64), L.astype(np.float64), 1)
            assert abs(row["fSlope"] - s) <= 1e-5 * abs(s), t
            assert abs(row["fIntercept"] - c) <= 1e-5 * max(1.0, abs(c), abs(s) * len(L)), t
            d = _gu_ranges(L, row["fSlope"], row["fIntercept"])
            assert row["nRangeLeft"] == int(max(0.0, float(d.max()))) + 1, t
            assert row["nRangeRight"] == int(max(0.0, float(-d.min()))) + 1, t
            assert np.all(d < row["nRangeLeft"]) and np.all(-d < row["nRangeRight"]), t
            assert row["nMagnification"] == 1 and row["fInterceptHigh"] == 0 and row["fInterceptLow"] == 0
            if len(L) > 2:
                r2 = np.corrcoef(np.arange(len(L)), L.astype(np.float64))[0, 1]
                assert abs(row["fRSquare"] - r2) <= 1e-5, t
            assert abs(row["fMultiple"] - np.sqrt(1 + np.float32(row["fSlope"]) ** 2)) <= 1e-6 * row["fMultiple"]
        b.save(str(tmp_path), "at")
        raw = np.fromfile(str(tmp_path / "at_regression_info"), dtype=ea.GU_REG_INFO)
        assert raw.tobytes() == g.tobytes()
    finally:
        b.close()


def test_gu_range_check_catches_a_short_window():
    """The containment check above fails on a window one position too
    narrow on either side (the slip of dropping the + 1), so it pins the
    ranges rather than restating them."""
    L = _gu_lists(21)[5]
    b = ea.build_index(*ea.lists_to_flat([L]), ea.RAW)
    try:
        row = b.gu_regression[0]
        d = _gu_ranges(L, row["fSlope"], row["fIntercept"])
        assert not np.all(d < row["nRangeLeft"] - 1)
        assert not np.all(-d < row["nRangeRight"] - 1)
    finally:
        b.close()
