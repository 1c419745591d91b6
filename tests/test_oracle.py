"""Pins the CPU oracle before it is trusted as the judge: reference golden
blocks, closed-form layouts, textbook least squares, brute-force intersection,
invariants of the codecs and the reference's checksum on hand-computed cases."""
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fixture_list():
    # src/gpu/Preprocess/ExtractList/data/InvertedList: 113,465 sorted docIDs (a GOV2 list)
    return np.fromfile(os.path.join(GOLD, "InvertedList.u32"), np.uint32)


def rand_list(rng, n, universe=1 << 24):
    return np.sort(rng.choice(universe, n, replace=False)).astype(np.uint32)


# ---------------------------------------------------------------- ParaPFor

def test_pfor_closed_form_no_exceptions():
    # 4 values of 3 bits packed LSB-first: 1 | 2<<3 | 3<<6 | 4<<9
    blk = oracle.pfor_encode([1, 2, 3, 4], 0.0)
    assert list(blk) == [3 << 26, 1 | 2 << 3 | 3 << 6 | 4 << 9]
    assert list(oracle.pfor_decode(blk, 4)[0]) == [1, 2, 3, 4]


def test_pfor_closed_form_exception():
    # n=10, FRAC 0.1 -> one exception allowed; lb = 1, position 9 (eb = 4), high 500 (hb = 9)
    v = [1] * 9 + [1000]
    blk = oracle.pfor_encode(v, 0.1)
    assert list(blk) == [(1 << 26) | (4 << 21) | (9 << 16) | 1, 0b0111111111, 9, 500]
    assert list(oracle.pfor_decode(blk, 10)[0]) == v


def test_pfor_straddling_values():
    # 17-bit values straddle word boundaries at 15, 30, ...
    v = np.arange(40, dtype=np.uint32) * 1000 + 65536  # max 104536: 17 bits
    blk = oracle.pfor_encode(v, 0.0)
    assert blk[0] >> 26 == 17 and len(blk) == 1 + (40 * 17 + 31) // 32
    assert np.array_equal(oracle.pfor_decode(blk, 40)[0], v)


def test_pfor_frac_budget_and_minimal_b():
    rng = np.random.default_rng(3)
    for frac in (0.0, 0.04, 0.1, 0.2):
        for _ in range(50):
            n = int(rng.integers(1, 257))
            v = rng.integers(0, 1 << int(rng.integers(1, 20)), n).astype(np.uint32)
            v[rng.choice(n, max(1, n // 8), replace=False)] = rng.integers(1 << 22, 1 << 28)
            b = oracle.pfor_best_b(v, frac)
            ex = int((v >= (1 << b)).sum())
            assert ex <= np.float32(frac) * np.float32(n)
            if b > 0:
                assert int((v >= (1 << (b - 1))).sum()) > np.float32(frac) * np.float32(n)
            blk = oracle.pfor_encode(v, frac)
            assert (blk[0] & 0xffff) == ex
            assert np.array_equal(oracle.pfor_decode(blk, n)[0], v)


def test_pfor_matches_reference_golden_blocks():
    g = np.load(os.path.join(GOLD, "parapfor_ref_blocks.npz"))
    for k in range(len(g["frac"])):
        v = g["values"][g["voff"][k]:g["voff"][k + 1]]
        ref_words = g["words"][g["woff"][k]:g["woff"][k + 1]]
        blk = oracle.pfor_encode(v, float(g["frac"][k]))
        assert np.array_equal(blk, ref_words), k
        dec, used = oracle.pfor_decode(ref_words, len(v))
        assert used == len(ref_words)
        assert np.array_equal(dec, g["decoded"][g["voff"][k]:g["voff"][k + 1]])
        assert np.array_equal(dec, v)


def test_pfor_matches_reference_library_fuzz():
    ref = oracle.ref_lib()
    if ref is None:
        pytest.skip("reference ParaPFor not built into oracle/_ref")
    rng = np.random.default_rng(9)
    for _ in range(400):
        n = int(rng.integers(1, 300))
        bits = int(rng.integers(0, 31))
        v = rng.integers(0, 1 << bits, n).astype(np.uint32) if bits else np.zeros(n, np.uint32)
        if rng.random() < 0.5:
            idx = rng.choice(n, int(rng.integers(0, n // 5 + 1)), replace=False)
            v[idx] = rng.integers(0, 1 << 30, len(idx)).astype(np.uint32)
        frac = float(rng.choice([0.0, 0.04, 0.1, 0.5]))
        out = np.zeros(4 * n + 8, np.uint32)
        w = ref.ref_pfor_encode(v.ctypes.data, n, frac, out.ctypes.data)
        assert np.array_equal(oracle.pfor_encode(v, frac), out[:w])


# ---------------------------------------------------------------- fits

def test_lrc_fit_is_least_squares():
    rng = np.random.default_rng(4)
    for n in (2, 3, 50, 1000, 100000):
        L = rand_list(rng, n)
        f = oracle.lrc_fit(L)
        slope, icpt = np.polyfit(np.arange(n, dtype=np.float64), L.astype(np.float64), 1)
        assert abs(float(f["fSlope"]) - slope) <= 1e-6 * abs(slope)
        assert abs(int(f["iIntercept"]) - int(icpt)) <= 1
        # search range covers every point: |(L_i - icpt)/slope - i| <= range
        px = (L.astype(np.float64) - icpt) / slope - np.arange(n)
        assert float(f["fSearchRange"]) >= (px.max() - min(px.min(), 0)) * (1 - 1e-6) - 1e-3


def test_seg_fit_exact_line():
    L = (5 + 7 * np.arange(128)).astype(np.uint32)
    f = oracle.seg_fit(L)
    assert float(f["fSlope"]) == 7.0 and float(f["fIntercept"]) == 5.0
    assert int(f["nRangeLeft"]) == 1 and int(f["nRangeRight"]) == 1
    ix = oracle.build_index([L], oracle.SEGLRC, seg=128)
    assert ix.words[0] >> 26 == 0  # all residuals zero -> lb 0
    assert ix.base[0] == 0


def test_fit_special_cases():
    f = oracle.seg_fit([123456])
    assert float(f["fSlope"]) == 1.0 and float(f["fIntercept"]) == 123456.0
    g = oracle.lrc_fit([99])
    assert float(g["fSlope"]) == 1.0 and int(g["iIntercept"]) == 99


# ---------------------------------------------------------------- codecs

CODEC_CASES = [(oracle.LRC, 256, 0.0, 0), (oracle.LRC, 128, 0.0, 0), (oracle.SEGLRC, 128, 0.0, 0),
               (oracle.SEGLRC, 256, 0.0, 0), (oracle.HS, 128, 0.0, 0), (oracle.HS, 128, 0.0, 13),
               (oracle.PFD, 128, 0.1, 0), (oracle.PFD, 256, 0.1, 0), (oracle.PFD, 128, 0.0, 0)]


@pytest.mark.parametrize("codec,seg,frac,shift", CODEC_CASES)
def test_roundtrip_fixture_and_synthetic(codec, seg, frac, shift):
    rng = np.random.default_rng(5)
    lists = [fixture_list(), rand_list(rng, 1), rand_list(rng, 129), rand_list(rng, 70000),
             np.arange(1000, 3000, dtype=np.uint32)]
    ix = oracle.build_index(lists, codec, seg=seg, frac=frac, hs_shift=shift)
    for t, L in enumerate(lists):
        assert np.array_equal(ix.decode_term(t), L)
    if codec in (oracle.LRC, oracle.SEGLRC, oracle.PFD):
        for t, L in enumerate(lists):
            s0, ns = int(ix.segnum[t]), int(ix.nsegs[t])
            assert np.array_equal(ix.seg_head[s0:s0 + ns], L[::seg])
            if codec == oracle.PFD:
                mids = [L[s * seg + min(seg, len(L) - s * seg) // 2] for s in range(ns)]
                assert np.array_equal(ix.seg_median[s0:s0 + ns], mids)


def test_residual_invariants():
    L = fixture_list()
    ix = oracle.build_index([L], oracle.SEGLRC, seg=128)
    # every segment's smallest residual is 0 and lb = bits(max residual)
    for s in range(int(ix.nsegs[0])):
        blk = ix.words[ix.seg_off[s]:]
        n = min(128, len(L) - 128 * s)
        vd, _ = oracle.pfor_decode(blk[:1 + n], n)
        assert vd.min() == 0
        assert blk[0] >> 26 == int(vd.max()).bit_length()
    lx = oracle.build_index([L], oracle.LRC, seg=256)
    allvd = []
    for s in range(int(lx.nsegs[0])):
        n = min(256, len(L) - 256 * s)
        allvd.append(oracle.pfor_decode(lx.words[lx.seg_off[s]:lx.seg_off[s] + 1 + n], n)[0])
    assert np.concatenate(allvd).min() == 0


def test_hs_table_closed_form():
    rng = np.random.default_rng(6)
    for n, shift in ((5000, 13), (300, 13), (100000, 0), (256, 0), (257, 0)):
        L = rand_list(rng, n, 1 << 25)
        ix = oracle.build_index([L], oracle.HS, hs_shift=shift)
        h = ix.hash_info
        if shift == 0:
            if n <= 256:
                assert list(h[:4]) == [int(L[-1]).bit_length(), 1, 0, n]
                continue
            nb_exp = (n + 255) // 256
            shift = int(L[-1]).bit_length() - nb_exp.bit_length()
        assert h[0] == shift
        nb = (int(L[-1]) >> shift) + 1
        assert h[1] == nb
        starts = np.searchsorted(L, (np.arange(nb + 1, dtype=np.int64) << shift), side="left")
        starts[-1] = n
        assert np.array_equal(h[2:3 + nb], starts)


# ---------------------------------------------------------------- queries

def test_eval_matches_bruteforce_and_special_cases():
    rng = np.random.default_rng(7)
    lists = [rand_list(rng, n) for n in (0, 1, 10, 500, 5000, 200000, 200000)]
    lists[6] = lists[5].copy()  # identical lists
    lists.append(np.arange(0, 100, dtype=np.uint32))           # disjoint from most
    lists.append(np.arange((1 << 24) - 100, 1 << 24, dtype=np.uint32))
    qs = [[0, 3], [1], [2, 5], [5, 6], [3, 4, 5], [7, 8], [4, 4], [], [6, 5, 4, 3, 2], [8, 5]]
    for codec, seg, frac, shift in CODEC_CASES:
        ix = oracle.build_index(lists, codec, seg=seg, frac=frac, hs_shift=shift)
        cum, res = oracle.eval_queries(ix, qs, nthreads=3)
        exp = [oracle.intersect_all([lists[t] for t in q]) if q else np.zeros(0, np.uint32) for q in qs]
        assert np.array_equal(cum, np.cumsum([len(e) for e in exp]))
        assert np.array_equal(res, np.concatenate(exp))
        cum2, res2 = oracle.eval_queries(ix, qs, full=True)
        assert np.array_equal(cum2, cum) and np.array_equal(res2, res)
        # special cases: empty list -> empty, identical -> the list, disjoint -> empty
        assert cum[0] == 0
        assert np.array_equal(exp[3], lists[5])
        assert len(exp[5]) == 0


def test_checksum_hand_computed():
    # queries with results [5,7,9], [], [4]: cs1 = 1 empty, cs2 = 4 results,
    # cs3 = (5^9^7) ^ (4^4^4) = 11 ^ 4 = 15
    assert oracle.checksum([3, 3, 4], [5, 7, 9, 4]) == (1, 4, 15)


def test_checksum_invariant_across_codecs():
    rng = np.random.default_rng(8)
    lists = [rand_list(rng, n) for n in (300, 3000, 30000, 300000)]
    qs = [[0, 1], [1, 2, 3], [0, 3], [2, 3], [3]]
    sums = set()
    for codec, seg, frac, shift in CODEC_CASES:
        ix = oracle.build_index(lists, codec, seg=seg, frac=frac, hs_shift=shift)
        sums.add(oracle.checksum(*oracle.eval_queries(ix, qs)))
    assert len(sums) == 1


def test_planted_mistake_is_caught():
    # a flipped residual bit in one block must break the round trip
    rng = np.random.default_rng(10)
    L = rand_list(rng, 1000)
    ix = oracle.build_index([L], oracle.SEGLRC, seg=128)
    ix.words[ix.seg_off[3] + 2] ^= 1 << 7
    assert not np.array_equal(ix.decode_term(0), L)


def test_c1_plain_query():
    """The plain-method configuration: one 2-term query, 1,000 and 1,000,000
    docIDs from [0, 2^24) (seed 1), LRC with 256-docID blocks; the oracle's
    decode-and-intersect equals numpy's set intersection."""
    from exclucene_amd import workloads as W
    a, b = W.c1_lists(1)
    ix = oracle.build_index([a, b], oracle.LRC, seg=256)
    cum, res = oracle.eval_queries(ix, [[0, 1], [1, 0]])
    want = np.intersect1d(a, b)
    assert np.array_equal(res[:cum[0]], want) and np.array_equal(res[cum[0]:], want)
    # expected size n_a * n_b / 2^24 = 59.6; a seeded draw lands near it
    assert 30 <= len(want) <= 95


def test_lrc_prediction_saturates_like_the_gpu():
    """float -> int32 of a fit prediction saturates (CUDA's cvt.rzi, the
    reference's GPU generator and decoder; gfx950's v_cvt_i32_f32), where a C
    cast on x86 would give INT32_MIN: for [0, 2^31 - 1] the slope is 2^31, so
    y(1) = (float)1 * slope = 2^31 must become 2^31 - 1."""
    top = (1 << 31) - 1
    o = oracle.build_index([np.array([0, top], np.uint32)], oracle.LRC, seg=256)
    slope = np.float32(o.lrc[0]["fSlope"])
    icpt = int(o.lrc[0]["iIntercept"])
    y1 = float(np.float32(1.0) * slope)
    assert y1 >= 2.0 ** 31
    w0 = int(o.uwlist[0]) + int(o.seg_off[0])
    vd, _ = oracle.pfor_decode(o.words[w0:w0 + 8], 2)
    sat = (int(vd[1]) + min(int(y1), top) + icpt) % (1 << 32)
    wrap = (int(vd[1]) - (1 << 31) + icpt) % (1 << 32)  # the x86 cast's INT32_MIN
    assert sat == top and wrap != top
    assert np.array_equal(o.decode_term(0), [0, top])


# ---------------------------------------------------------------- fits off an exact line

def test_seg_fit_is_least_squares():
    """SegLRC segment and list fits on random lists against numpy's least
    squares: slope within float rounding of polyfit's, intercept within the
    float rounding of ya - slope * xa, and the recorded ranges cover every
    element's position from the fit's own float prediction."""
    rng = np.random.default_rng(41)
    fx = fixture_list()
    cases = [rand_list(rng, n, 1 << 25) for n in (2, 3, 17, 128, 256, 4096, 100000)]
    cases += [fx[:128], fx[5000:5256], fx]
    for L in cases:
        n = len(L)
        f = oracle.seg_fit(L)
        i = np.arange(n, dtype=np.float64)
        slope, icpt = np.polyfit(i, L.astype(np.float64), 1)
        fs, fi = float(f["fSlope"]), float(f["fIntercept"])
        assert abs(fs - slope) <= 2e-7 * abs(slope) + 1e-12
        # intercept = mean(L) - float(slope) * mean(i), rounded to float
        ref_i = L.astype(np.float64).mean() - fs * i.mean()
        assert abs(fi - ref_i) <= abs(ref_i) * 2 ** -23 + 1e-6
        assert abs(fi - icpt) <= abs(slope - fs) * i.mean() + abs(icpt) * 2 ** -23 + 1e-6
        # coverage: every element lies inside [pred - nRangeLeft, pred + nRangeRight]
        px = (L.astype(np.float32) - np.float32(fi)) / np.float32(fs)
        dev = px.astype(np.float64) - i
        assert (dev <= int(f["nRangeLeft"])).all() and (-dev <= int(f["nRangeRight"])).all()
        # and the ranges are the tightest such integers (+1 as the reference records them)
        assert int(f["nRangeLeft"]) == int(max(dev.max(), 0.0)) + 1
        assert int(f["nRangeRight"]) == int(max(-dev.min(), 0.0)) + 1


def test_seg_fit_planted_slope_slip_fails_polyfit():
    """A plausible slip (the slope's denominator off by one term) would move
    the slope off polyfit's by far more than the check's tolerance."""
    rng = np.random.default_rng(42)
    L = rand_list(rng, 128, 1 << 25).astype(np.float64)
    i = np.arange(128, dtype=np.float64)
    xa, ya = i.mean(), L.mean()
    good = ((i - xa) * (L - ya)).sum() / ((i - xa) ** 2).sum()
    slip = ((i - xa) * (L - ya)).sum() / ((i - xa) ** 2).sum()[()] / (1 + 1 / 127)
    slope = np.polyfit(i, L, 1)[0]
    assert abs(good - slope) <= 2e-7 * abs(slope)
    assert abs(slip - slope) > 2e-7 * abs(slope)


# ---------------------------------------------------------------- learned models

def test_pla_knots_are_the_empirical_cdf():
    """PLA-256 knots equal np.searchsorted of the knot positions in the
    fp32-scaled docIDs (elements strictly below each knot), with x0 = first
    docID and xs = 256 / (span + 1) in fp32."""
    rng = np.random.default_rng(43)
    lists = [rand_list(rng, n, 1 << 25) for n in (1, 2, 300, 70000)]
    lists += [fixture_list(), (np.arange(4096, dtype=np.uint32) * 7 + 11)]
    for L in lists:
        x0, xs, y = oracle.pla_fit(L)
        assert np.float32(x0) == np.float32(L[0])
        assert np.float32(xs) == np.float32(256.0) / (np.float32(L[-1]) - np.float32(L[0]) + np.float32(1.0))
        t = (L.astype(np.float32) - np.float32(x0)) * np.float32(xs)
        assert (np.diff(t) >= 0).all()
        want = np.searchsorted(t, np.arange(257, dtype=np.float32), side="left")
        assert np.array_equal(y.astype(np.int64), want)
        # (the last docID's scaled value can round up to 256 in fp32, so
        # y[256] counts the elements below it, not always all of them)
        assert y[0] == 0 and y[256] >= len(L) - 1
    # an exactly uniform list: knot p holds ceil(p * n / 256) elements
    L = np.arange(0, 256 * 64, 64, dtype=np.uint32)
    _, _, y = oracle.pla_fit(L)
    assert np.array_equal(y[:256], np.arange(256))


def test_pla_prediction_interpolates_the_knots():
    """The PLA model's prediction is the piecewise-linear interpolation of
    its knots (np.interp), so its recorded max error is the exhaustive max of
    |interp - position| up to fp32 rounding."""
    rng = np.random.default_rng(44)
    L = rand_list(rng, 50000, 1 << 25)
    x0, xs, y = oracle.pla_fit(L)
    t = np.clip((L.astype(np.float32) - np.float32(x0)) * np.float32(xs), 0, 256).astype(np.float64)
    pred = np.interp(t, np.arange(257, dtype=np.float64), y.astype(np.float64))
    err = np.abs(pred - np.arange(len(L))).max()
    got = oracle.model_max_err(1, y, x0, xs, L)
    assert abs(got - err) <= 1e-3 * max(1.0, err)


def test_plaq_knots_are_list_quantiles():
    """PLA-Q knots sit at positions floor(j (n - 1) / 256): x = the fp32
    docID there, y = the position, each slope the rise over the run of its
    piece (0 where the fp32 knots coincide)."""
    rng = np.random.default_rng(45)
    for L in [rand_list(rng, n, 1 << 25) for n in (2, 300, 70000)] + [fixture_list()]:
        prm = oracle.plaq_fit(L)
        q = (np.arange(257, dtype=np.uint64) * np.uint64(len(L) - 1)) // np.uint64(256)
        x, y, s = prm[:257], prm[257:514], prm[514:770]
        assert np.array_equal(x, L[q].astype(np.float32))
        assert np.array_equal(y, q.astype(np.float32))
        dx = np.diff(x)
        want = np.where(dx > 0, np.diff(y) / np.where(dx > 0, dx, np.float32(1)), np.float32(0)).astype(np.float32)
        assert np.array_equal(s, want)


def test_plaq_prediction_and_error_bound():
    """PLA-Q predicts each knot's own position at its docID and the linear
    interpolation of the knots in between (np.interp, up to fp32 rounding);
    its max error over any list stays within one piece's positions plus the
    fp32 spacing of the docIDs, here on a clustered list a line fits poorly."""
    from exclucene_amd import workloads as W
    rng = np.random.default_rng(46)
    centres = rng.integers(0, 1 << 25, 64)
    halfw = (2.0 ** rng.uniform(-1, 3, 64) * (1 << 14)).astype(np.int64)
    for L in (rand_list(rng, 50000, 1 << 25), W.clustered_list(rng, 120000, centres, halfw)):
        prm = oracle.plaq_fit(L)
        x, y = prm[:257], prm[257:514]
        for j in range(0, 257, 16):
            assert oracle.plaq_pred(prm, int(L[int(y[j])])) == y[j]
        vs = np.sort(rng.choice(L, 200, replace=False))
        got = np.array([oracle.plaq_pred(prm, int(v)) for v in vs])
        want = np.interp(vs.astype(np.float32).astype(np.float64), x.astype(np.float64), y.astype(np.float64))
        assert np.abs(got - want).max() <= 1e-3 * len(L)
        err = oracle.model_max_err(3, prm, 0.0, 0.0, L)
        assert err <= (len(L) - 1) / 256 + 2 + 4, err
    # one line misses the clustered list by far more
    f = np.polyfit(np.arange(len(L)), L.astype(np.float64), 1)
    line_err = np.abs((L - f[1]) / f[0] - np.arange(len(L))).max()
    assert line_err > 8 * err


def _fd_max_rel_err(params, L, samples, grad_fn, h=1e-6):
    """Largest relative difference between grad_fn's gradient and float64
    central finite differences of the oracle's loss."""
    _, g = grad_fn(params)
    fd = np.zeros_like(g)
    for k in range(len(params)):
        p1 = params.copy()
        p1[k] += h
        p0 = params.copy()
        p0[k] -= h
        fd[k] = (oracle.mlp_loss_grad(p1, L, samples)[0] - oracle.mlp_loss_grad(p0, L, samples)[0]) / (2 * h)
    scale = np.maximum(np.abs(fd), 1e-3 * np.abs(fd).max())
    return float((np.abs(g - fd) / scale).max())


def _mlp_params(rng):
    # weights drawn so that every hidden unit is active on part of [0, 1]
    w1 = rng.normal(0.0, 2.0, 32)
    b1 = rng.uniform(-1.0, 1.0, 32)
    w2 = rng.normal(0.0, 0.5, 32)
    return np.concatenate([w1, b1, w2, [rng.normal(0.0, 0.1)]]).astype(np.float64)


def test_mlp_gradient_matches_finite_differences():
    """oracle.mlp_loss_grad's analytic gradient (the judge of the GPU
    training kernel's backward pass) equals float64 central differences of
    its own loss at random parameters."""
    rng = np.random.default_rng(45)
    for n, S in ((5000, 256), (70000, 1024), (300, 300)):
        L = rand_list(rng, n, 1 << 25)
        for _ in range(2):
            params = _mlp_params(rng)
            err = _fd_max_rel_err(params, L, S, lambda p: oracle.mlp_loss_grad(p, L, S))
            assert err < 1e-5, err


@pytest.mark.parametrize("slip", ["relu_mask", "scale"])
def test_mlp_gradient_planted_slip_is_caught(slip):
    """A gradient with the ReLU mask dropped, or without the 2/S factor,
    fails the same finite-difference check."""
    rng = np.random.default_rng(46)
    L = rand_list(rng, 5000, 1 << 25)
    params = _mlp_params(rng)
    err = _fd_max_rel_err(params, L, 256, lambda p: oracle.mlp_loss_grad(p, L, 256, _plant=slip))
    assert err > 1e-2, err


# ---------------------------------------------------------------- uncompressed hash segments

def _int_ceil_log2(x):
    """ceil(log2 x) in exact integer arithmetic (0 for x <= 1)."""
    return (int(x) - 1).bit_length() if x > 1 else 0


def test_gu_hashseg_table_closed_form():
    """The uncompressed hash-segment tables (cuda_hashSegment.cu:815-947)
    against their closed form: lists of at most 256 docIDs get
    [ceil(log2 max), 0, n]; longer lists [shift][start_0 .. start_nb] with
    shift = ceil(log2 max) - ceil(log2(n / 256)), nb = (max >> shift) + 1 and
    start_b = the number of docIDs below b << shift.  The reference takes the
    logarithms in doubles, which round above the integer for max = 2^29 and
    2^31 only: there its bit count, and so its shift, is one more."""
    rng = np.random.default_rng(47)
    lists = [rand_list(rng, n, 1 << 25) for n in (1, 2, 255, 256, 257, 511, 512, 513, 4096, 70000, 1 << 20)]
    lists += [fixture_list(), np.arange(1000, 1000 + 777, dtype=np.uint32)]
    lists += [np.array([5, 1 << k], np.uint32) for k in (10, 24, 28, 29, 30)]
    lists += [np.sort(np.append(rng.choice(1 << 20, 600, replace=False), 1 << 29)).astype(np.uint32)]
    for L in lists:
        n, mx = len(L), int(L[-1])
        bits = _int_ceil_log2(mx) + (1 if mx in (1 << 29, 1 << 31) else 0)
        h = oracle.gu_hashseg_table(L)
        if n <= 256:
            assert list(h) == [bits, 0, n], (n, mx, list(h))
            continue
        # ceil(log2(n / 256)) = ceil(log2 n) - 8 for n > 256
        shift = bits - (_int_ceil_log2(n) - 8)
        assert h[0] == shift, (n, mx, h[0], shift)
        nb = (mx >> shift) + 1
        assert len(h) == nb + 2
        starts = np.searchsorted(L, np.arange(nb, dtype=np.int64) << shift, side="left")
        assert np.array_equal(h[1:1 + nb], starts) and h[1 + nb] == n
    # a fixed shift: the same walk with that shift, for short lists too
    for L in lists[:4]:
        h = oracle.gu_hashseg_table(L, shift=13)
        nb = (int(L[-1]) >> 13) + 1
        assert h[0] == 13 and len(h) == nb + 2 and h[-1] == len(L)
        assert np.array_equal(h[1:1 + nb], np.searchsorted(L, np.arange(nb, dtype=np.int64) << 13, side="left"))
    assert list(oracle.gu_hashseg_table(np.zeros(0, np.uint32))) == [0, 0, 0]
