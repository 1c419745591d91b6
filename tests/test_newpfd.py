"""NewPFD (d-gaps in NewPFor blocks with Simple16-coded exceptions,
src/gpu/Compressed/src/NewPFD): the oracle against the reference's own
NewPFor.cpp (golden blocks in tests/golden/newpfor_ref_blocks.npz, made by
tests/golden/make_golden.py from oracle/_ref), Simple16 against its closed
form, whole-list round trips, and the library's host encoder byte-identical
to the oracle's."""
import os

import numpy as np
import pytest

import exclucene_amd as ea
import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "newpfor_ref_blocks.npz")
# Simple16 layouts as (count, width) runs, from the paper's table (Zhang, Long, Suel 2008)
LAYOUTS = [[(28, 1)], [(7, 2), (14, 1)], [(7, 1), (7, 2), (7, 1)], [(14, 1), (7, 2)], [(14, 2)],
           [(1, 4), (8, 3)], [(1, 3), (4, 4), (3, 3)], [(7, 4)], [(4, 5), (2, 4)], [(2, 4), (4, 5)],
           [(3, 6), (2, 5)], [(2, 5), (3, 6)], [(4, 7)], [(1, 10), (2, 9)], [(2, 14)], [(1, 28)]]


def widths(k):
    return [w for c, w in LAYOUTS[k] for _ in range(c)]


def test_golden_blocks_match_reference():
    z = np.load(GOLD)
    vals, words, woff, voff, frac, dec = z["values"], z["words"], z["woff"], z["voff"], z["frac"], z["decoded"]
    nex = 0
    for i in range(len(frac)):
        v = vals[voff[i]:voff[i + 1]]
        w = words[woff[i]:woff[i + 1]]
        assert np.array_equal(dec[voff[i]:voff[i + 1]], v), i  # the reference round-trips its own blocks
        mine = oracle.npfor_encode(v, float(frac[i]))
        assert np.array_equal(mine, w), i
        out, used = oracle.npfor_decode(w, len(v))
        assert used == len(w) and np.array_equal(out, v), i
        nex += int(w[0] & 0xFFFF) > 0
    assert nex >= 60


def test_simple16_closed_form():
    rng = np.random.default_rng(3)
    for k in range(16):
        ws = widths(k)
        assert sum(ws) <= 28 and len(ws) == [28, 21, 21, 21, 14, 9, 8, 7, 6, 6, 5, 5, 4, 3, 2, 1][k]
    for trial in range(300):
        n = int(rng.integers(1, 90))
        v = (rng.integers(0, 1 << 28, n) >> rng.integers(0, 28, n)).astype(np.uint32)
        enc = oracle.s16_encode(v)
        # decode by the closed form, checking that every word's selector is
        # the first one whose slots hold the values it carries
        got, i = [], 0
        for word in enc:
            k = int(word) >> 28
            ws = widths(k)
            vol = min(len(ws), n - i)
            sh = 0
            for j in range(len(ws)):
                x = (int(word) >> sh) & ((1 << ws[j]) - 1)
                sh += ws[j]
                if j < vol:
                    got.append(x)
                else:
                    assert x == 0
            for k2 in range(k):
                w2 = widths(k2)
                vol2 = min(len(w2), n - i)
                assert any(int(v[i + j]) >> w2[j] for j in range(vol2)), (trial, k, k2)
            i += vol
        assert i == n and np.array_equal(np.array(got, np.uint32), v)
        d, used = oracle.s16_decode(enc, n)
        assert used == len(enc) and np.array_equal(d, v)
    with pytest.raises(ValueError):
        oracle.s16_encode(np.array([1, 1 << 28], np.uint32))


def test_high_parts_beyond_28_bits_raise_b():
    v = np.array([1 << 30, 1, 2, 3, 1, 2, 3, 1], np.uint32)
    w = oracle.npfor_encode(v, 0.5)
    assert (int(w[0]) >> 26) >= 3  # (1 << 30) >> b fits 28 bits
    out, _ = oracle.npfor_decode(w, len(v))
    assert np.array_equal(out, v)


def _lists(seed):
    rng = np.random.default_rng(seed)
    lists = [np.array([5], np.uint32), np.array([0, 1, 2, 3], np.uint32), np.arange(0, 256 * 9, 9, dtype=np.uint32)]
    for n in (255, 256, 257, 1000, 4097, 60000):
        lists.append(np.sort(rng.choice(1 << 24, n, replace=False)).astype(np.uint32))
    # clustered docIDs with rare long jumps: exceptions in most blocks
    g = rng.integers(1, 8, 20000).astype(np.uint64)
    g[rng.choice(20000, 400, replace=False)] = rng.integers(1 << 12, 1 << 20, 400)
    lists.append(np.cumsum(g).astype(np.uint32))
    return lists


@pytest.mark.parametrize("seg", [128, 256])
@pytest.mark.parametrize("frac", [0.0, 0.1, 0.32])
def test_lists_round_trip_and_builder_matches_oracle(seg, frac):
    lists = _lists(9)
    o = oracle.build_index(lists, oracle.NPFD, seg=seg, frac=frac)
    for t, L in enumerate(lists):
        assert np.array_equal(o.decode_term(t), L), t
    docs, off = ea.lists_to_flat(lists)
    b = ea.build_index(docs, off, ea.NEWPFD, seg=seg, frac=frac)
    try:
        assert np.array_equal(b.ind2, o.words[:o.total_words])
        assert np.array_equal(b.seg_offset, o.seg_off) and np.array_equal(b.seg_head, o.seg_head)
        assert np.array_equal(b.seg_median, o.seg_median) and np.array_equal(b.seg_num, o.segnum)
        assert np.array_equal(b.ind1["m_off"], o.uwlist * 4)
    finally:
        b.close()


def test_reference_fixture_round_trip():
    L = np.fromfile(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "InvertedList.u32"),
                    np.uint32)
    for frac in (0.0, 0.1, 0.64):
        o = oracle.build_index([L], oracle.NPFD, seg=256, frac=frac)
        assert np.array_equal(o.decode_term(0), L)


def test_planted_mistake_is_caught():
    """Dropping the + 1 of the position deltas (a plausible restatement slip)
    no longer reproduces the reference's values: the golden check sees it."""
    z = np.load(GOLD)
    i = next(i for i in range(len(z["frac"])) if int(z["words"][z["woff"][i]]) & 0xFFFF > 1)
    w = z["words"][z["woff"][i]:z["woff"][i + 1]].copy()
    v = z["values"][z["voff"][i]:z["voff"][i + 1]]
    n, b, en = len(v), int(w[0]) >> 26, int(w[0]) & 0xFFFF
    lw = w.copy()
    lw[0] = b << 26  # the same block without its exceptions: low parts only
    low, _ = oracle.npfor_decode(lw, n)
    allv, _ = oracle.s16_decode(w[1 + (n * b + 31) // 32:], 2 * en)

    def patched(plus):
        out = low.astype(np.uint64)
        p = int(allv[0])
        for m in range(en):
            if p < n:
                out[p] += int(allv[en + m]) << b
            if m + 1 < en:
                p += int(allv[m + 1]) + plus
        return out

    assert np.array_equal(patched(1), v) and not np.array_equal(patched(0), v)
