"""CPU oracle for the compressed posting-list decode + intersect path.

TEST INFRASTRUCTURE ONLY.  This package is the checker: tests/, the smoke()
entry and the cpu_baseline leg of bench.py use it; the exclucene_amd product
never imports it.  The numerics live in oracle.c (a C restatement of the
reference's ParaPFor / LRC / SegLRC / HS_SegLRC / ParaPFD / NewPFD codecs and of its
batched intersection semantics, with file:line citations); this module only
binds it with ctypes and assembles whole indexes from per-list encoders the
way the reference's *_Compression programs lay them out
(src/gpu/Compressed/src/LRC/LRC_Compression/main.cpp:163-216).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libref_parapfor.so")
REF_NPFOR_PATH = os.path.join(HERE, "_ref", "libref_newpfor.so")

LRC, SEGLRC, HS, PFD, NPFD = 0, 1, 2, 3, 4
CODEC_NAMES = {LRC: "LRC", SEGLRC: "SegLRC", HS: "HS_SegLRC", PFD: "ParaPFD"}

LRC_INFO = np.dtype([("fSlope", "<f4"), ("iIntercept", "<i4"), ("fSearchRange", "<f4")])
SEG_INFO = np.dtype([("fSlope", "<f4"), ("fIntercept", "<f4"),
                     ("nRangeLeft", "<u4"), ("nRangeRight", "<u4")])

_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p


class OrcIndex(ctypes.Structure):
    _fields_ = [("codec", ctypes.c_int32), ("seg", ctypes.c_uint32),
                ("nterms", ctypes.c_uint32), ("pad_", ctypes.c_uint32),
                ("total_words", ctypes.c_uint64),
                ("ucnt", _vp), ("uwlist", _vp), ("segnum", _vp), ("words", _vp),
                ("seg_off", _vp), ("seg_head", _vp), ("seg_median", _vp),
                ("lrc", _vp), ("fit", _vp), ("base", _vp),
                ("hash_off", _vp), ("hash_info", _vp)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_pfor_best_b.restype = ctypes.c_uint32
        L.orc_pfor_best_b.argtypes = [_vp, ctypes.c_uint32, ctypes.c_float]
        L.orc_pfor_encode.restype = ctypes.c_uint32
        L.orc_pfor_encode.argtypes = [_vp, ctypes.c_uint32, ctypes.c_float, _vp]
        L.orc_pfor_decode.restype = ctypes.c_uint32
        L.orc_pfor_decode.argtypes = [_vp, ctypes.c_uint32, _vp]
        L.orc_lrc_fit.argtypes = [_vp, ctypes.c_uint32, _vp]
        L.orc_seg_fit.argtypes = [_vp, ctypes.c_uint32, _vp]
        L.orc_lrc_encode_list.restype = ctypes.c_uint64
        L.orc_lrc_encode_list.argtypes = [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float,
                                          _vp, _vp, _vp, _vp]
        L.orc_seglrc_encode_list.restype = ctypes.c_uint64
        L.orc_seglrc_encode_list.argtypes = [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float,
                                             _vp, _vp, _vp, _vp, _vp, _vp]
        L.orc_hs_table.restype = ctypes.c_uint32
        L.orc_hs_table.argtypes = [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _vp]
        L.orc_gu_hashseg_table.restype = ctypes.c_uint32
        L.orc_gu_hashseg_table.argtypes = [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _vp]
        L.orc_hs_encode_list.restype = ctypes.c_uint64
        L.orc_hs_encode_list.argtypes = [_vp, ctypes.c_uint32, _vp, ctypes.c_float, _vp, _vp, _vp, _vp]
        L.orc_pfd_encode_list.restype = ctypes.c_uint64
        L.orc_pfd_encode_list.argtypes = [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float,
                                          _vp, _vp, _vp, _vp]
        for f in ("orc_s16_encode", "orc_s16_decode", "orc_npfor_decode"):
            getattr(L, f).restype = ctypes.c_uint32
            getattr(L, f).argtypes = [_vp, ctypes.c_uint32, _vp]
        L.orc_npfor_encode.restype = ctypes.c_uint32
        L.orc_npfor_encode.argtypes = [_vp, ctypes.c_uint32, ctypes.c_float, _vp]
        L.orc_npfd_encode_list.restype = ctypes.c_uint64
        L.orc_npfd_encode_list.argtypes = [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float,
                                           _vp, _vp, _vp, _vp]
        L.orc_decode_term.restype = ctypes.c_int
        L.orc_decode_term.argtypes = [ctypes.POINTER(OrcIndex), ctypes.c_uint32, _vp]
        L.orc_checksum.argtypes = [_vp, _vp, ctypes.c_uint32, _vp]
        L.orc_eval_queries.restype = ctypes.c_int
        L.orc_eval_queries.argtypes = [ctypes.POINTER(OrcIndex), _vp, _vp, ctypes.c_uint32, ctypes.c_int,
                                       _vp, _vp, ctypes.c_uint64, _vp]
        L.orc_eval_queries_full.restype = ctypes.c_int
        L.orc_eval_queries_full.argtypes = [ctypes.POINTER(OrcIndex), _vp, _vp, ctypes.c_uint32,
                                            _vp, _vp, ctypes.c_uint64]
        L.orc_intersect2.restype = ctypes.c_uint64
        L.orc_intersect2.argtypes = [_vp, ctypes.c_uint64, _vp, ctypes.c_uint64, _vp]
        L.orc_pla_fit.argtypes = [_vp, ctypes.c_uint32, _vp, _vp, _vp]
        L.orc_pla_pred.restype = ctypes.c_float
        L.orc_pla_pred.argtypes = [_vp, ctypes.c_float, ctypes.c_float, ctypes.c_uint32]
        L.orc_mlp_pred.restype = ctypes.c_float
        L.orc_mlp_pred.argtypes = [_vp, ctypes.c_float, ctypes.c_float, ctypes.c_uint32, ctypes.c_uint32]
        L.orc_plaq_fit.argtypes = [_vp, ctypes.c_uint32, _vp]
        L.orc_plaq_pred.restype = ctypes.c_float
        L.orc_plaq_pred.argtypes = [_vp, ctypes.c_uint32]
        L.orc_model_max_err.restype = ctypes.c_float
        L.orc_model_max_err.argtypes = [ctypes.c_int, _vp, ctypes.c_float, ctypes.c_float, _vp, ctypes.c_uint32]
        L.orc_query_bytes.restype = ctypes.c_uint64
        L.orc_query_bytes.argtypes = [ctypes.POINTER(OrcIndex), _vp, _vp, ctypes.c_uint32, _vp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data if a is not None and a.size else None


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


# ---------------------------------------------------------------- block codec

def pfor_encode(vals, frac=0.0):
    v = _u32(vals)
    out = np.zeros(4 * len(v) + 8, np.uint32)
    w = lib().orc_pfor_encode(_p(v), len(v), frac, _p(out))
    if w == 0:
        raise ValueError("unrepresentable ParaPFor block")
    return out[:w].copy()


def pfor_decode(words, n):
    w = np.zeros(len(words) + 2, np.uint32)
    w[:len(words)] = words
    out = np.zeros(max(n, 1), np.uint32)
    used = lib().orc_pfor_decode(_p(w), n, _p(out))
    return out[:n], used


def pfor_best_b(vals, frac=0.0):
    v = _u32(vals)
    return lib().orc_pfor_best_b(_p(v), len(v), frac)


def s16_encode(vals):
    v = _u32(vals)
    out = np.zeros(len(v) + 2, np.uint32)
    w = lib().orc_s16_encode(_p(v), len(v), _p(out))
    if len(v) and w == 0:
        raise ValueError("a value needs more than 28 bits")
    return out[:w].copy()


def s16_decode(words, n):
    w = np.zeros(len(words) + 1, np.uint32)
    w[:len(words)] = words
    out = np.zeros(n + 28, np.uint32)
    used = lib().orc_s16_decode(_p(w), n, _p(out))
    return out[:n], used


def npfor_encode(vals, frac=0.0):
    v = _u32(vals)
    out = np.zeros(4 * len(v) + 8, np.uint32)
    w = lib().orc_npfor_encode(_p(v), len(v), frac, _p(out))
    if w == 0:
        raise ValueError("unrepresentable NewPFor block")
    return out[:w].copy()


def npfor_decode(words, n):
    w = np.zeros(len(words) + 2, np.uint32)
    w[:len(words)] = words
    out = np.zeros(max(n, 1), np.uint32)
    used = lib().orc_npfor_decode(_p(w), n, _p(out))
    return out[:n], used


def ref_npfor_lib():
    """The reference's own NewPFor.cpp (+ Simple16.h), compiled into _ref/
    (None if not built)."""
    if not os.path.exists(REF_NPFOR_PATH):
        return None
    L = ctypes.CDLL(REF_NPFOR_PATH)
    L.ref_npfor_encode.restype = ctypes.c_uint32
    L.ref_npfor_encode.argtypes = [_vp, ctypes.c_uint32, ctypes.c_float, _vp]
    L.ref_npfor_decode.restype = ctypes.c_uint32
    L.ref_npfor_decode.argtypes = [_vp, ctypes.c_uint32, _vp]
    return L


def ref_lib():
    """The reference's own ParaPFor.cpp, compiled into _ref/ (None if not built)."""
    if not os.path.exists(REF_LIB_PATH):
        return None
    L = ctypes.CDLL(REF_LIB_PATH)
    L.ref_pfor_encode.restype = ctypes.c_uint32
    L.ref_pfor_encode.argtypes = [_vp, ctypes.c_uint32, ctypes.c_float, _vp]
    L.ref_pfor_decode.restype = ctypes.c_uint32
    L.ref_pfor_decode.argtypes = [_vp, ctypes.c_uint32, _vp]
    return L


def gu_hashseg_table(L, target=256, shift=0):
    """The uncompressed hash-segment table of sorted list L:
    [shift][start_0 .. start_nb] (oracle.c: orc_gu_hashseg_table)."""
    a = _u32(L)
    if shift:
        words = (int(a[-1]) >> shift) + 3 if len(a) else 3
    else:  # nb <= 2^ceil(log2(n / target)) + 1 (one more where the double log rounds up)
        words = 4 * len(a) // max(target, 1) + 8
    out = np.zeros(words + 3, np.uint32)
    k = lib().orc_gu_hashseg_table(_p(a) if len(a) else None, len(a), target, shift, out.ctypes.data)
    return out[:k].copy()


def lrc_fit(L):
    a = _u32(L)
    info = np.zeros(1, LRC_INFO)
    lib().orc_lrc_fit(_p(a), len(a), info.ctypes.data)
    return info[0]


def seg_fit(L):
    a = _u32(L)
    info = np.zeros(1, SEG_INFO)
    lib().orc_seg_fit(_p(a), len(a), info.ctypes.data)
    return info[0]


# ---------------------------------------------------------------- whole index

class Index:
    """A compressed index in the reference's layout (arrays of the .ind1/.ind2/
    .segNum/.segOffset/.segHead/.segMedian/.regression_info/.ind2_distance_base/
    .HashSeg_offset/.HashSeg_info files), plus the oracle struct over them."""

    def __init__(self, **kw):
        self.__dict__.update(kw)
        self._struct = None

    def struct(self):
        if self._struct is None:
            s = OrcIndex()
            s.codec, s.seg, s.nterms = self.codec, self.seg, len(self.ucnt)
            s.total_words = self.total_words
            for name in ("ucnt", "uwlist", "segnum", "words", "seg_off", "seg_head", "seg_median",
                         "lrc", "fit", "base", "hash_off", "hash_info"):
                arr = getattr(self, name, None)
                setattr(s, name, _p(arr) if arr is not None else None)
            self._struct = s
        return self._struct

    def decode_term(self, t):
        n = int(self.ucnt[t])
        out = np.zeros(max(n, 1), np.uint32)
        lib().orc_decode_term(ctypes.byref(self.struct()), t, _p(out))
        return out[:n]

    def compressed_bytes(self):
        tot = self.total_words * 4 + len(self.ucnt) * (16 + 4)
        for name in ("seg_off", "seg_head", "seg_median", "lrc", "fit", "base", "hash_off", "hash_info"):
            arr = getattr(self, name, None)
            if arr is not None:
                tot += arr.nbytes
        return tot


PAD_WORDS = 4


def build_index(lists, codec, seg=128, frac=0.0, hs_target=256, hs_shift=0):
    """Encode `lists` (sorted, distinct uint32 docIDs each) into one index."""
    L_ = lib()
    lists = [_u32(x) for x in lists]
    nt = len(lists)
    ucnt = np.array([len(x) for x in lists], np.uint32)
    if codec == HS:
        tabs = []
        for x in lists:
            if hs_shift:
                nb_bound = (int(x[-1]) >> hs_shift) + 1 if len(x) else 1
            else:
                nb_bound = 2 * ((len(x) + hs_target - 1) // hs_target) + 2
            info = np.zeros(nb_bound + 4, np.uint32)
            k = L_.orc_hs_table(_p(x) if len(x) else None, len(x), hs_target, hs_shift, _p(info))
            tabs.append(info[:k].copy())
        nsegs = np.array([int(t[1]) for t in tabs], np.int64)
    else:
        nsegs = np.array([(len(x) + seg - 1) // seg for x in lists], np.int64)
    segnum = np.zeros(nt, np.uint32)
    segnum[1:] = np.cumsum(nsegs)[:-1]
    S = int(nsegs.sum())
    cap = int(sum(3 * len(x) + 4 * int(ns) + 8 for x, ns in zip(lists, nsegs)))
    words = np.zeros(cap + PAD_WORDS, np.uint32)
    uwlist = np.zeros(nt, np.uint64)
    seg_off = np.zeros(max(S, 1), np.uint32)
    seg_head = seg_median = lrc = fit = base = hash_off = hash_info = None
    if codec in (LRC, SEGLRC, PFD, NPFD):
        seg_head = np.zeros(max(S, 1), np.uint32)
    if codec in (PFD, NPFD):
        seg_median = np.zeros(max(S, 1), np.uint32)
    if codec == LRC:
        lrc = np.zeros(nt, LRC_INFO)
    if codec in (SEGLRC, HS):
        fit = np.zeros(max(S, 1), SEG_INFO)
        base = np.zeros(max(S, 1), np.int32)
    list_fit = np.zeros(nt, SEG_INFO) if codec == SEGLRC else None
    if codec == HS:
        hash_off = np.zeros(nt, np.uint32)
        hash_off[1:] = np.cumsum([len(t) for t in tabs])[:-1]
        hash_info = np.concatenate(tabs) if tabs else np.zeros(1, np.uint32)
    w = 0
    for t, x in enumerate(lists):
        uwlist[t] = w
        sn = int(segnum[t])
        n = len(x)
        wp = words[w:].ctypes.data
        if codec == LRC:
            k = L_.orc_lrc_encode_list(_p(x), n, seg, frac, lrc[t:].ctypes.data, wp,
                                       seg_head[sn:].ctypes.data, seg_off[sn:].ctypes.data)
        elif codec == SEGLRC:
            k = L_.orc_seglrc_encode_list(_p(x), n, seg, frac, fit[sn:].ctypes.data, base[sn:].ctypes.data,
                                          wp, seg_head[sn:].ctypes.data, seg_off[sn:].ctypes.data,
                                          list_fit[t:].ctypes.data)
        elif codec == HS:
            k = L_.orc_hs_encode_list(_p(x), n, _p(tabs[t]), frac, fit[sn:].ctypes.data,
                                      base[sn:].ctypes.data, wp, seg_off[sn:].ctypes.data)
        elif codec in (PFD, NPFD):
            enc = L_.orc_pfd_encode_list if codec == PFD else L_.orc_npfd_encode_list
            k = enc(_p(x), n, seg, frac, wp, seg_head[sn:].ctypes.data,
                    seg_median[sn:].ctypes.data, seg_off[sn:].ctypes.data)
        else:
            raise ValueError(codec)
        if n and k == 0:
            raise ValueError("unrepresentable block in list %d" % t)
        w += int(k)
    words = words[:w + PAD_WORDS].copy()
    return Index(codec=codec, seg=seg, frac=frac, ucnt=ucnt, uwlist=uwlist, segnum=segnum,
                 words=words, total_words=w, seg_off=seg_off, seg_head=seg_head,
                 seg_median=seg_median, lrc=lrc, fit=fit, base=base, list_fit=list_fit,
                 hash_off=hash_off, hash_info=hash_info, nsegs=nsegs)


def _queries_flat(queries):
    qoff = np.zeros(len(queries) + 1, np.uint64)
    qoff[1:] = np.cumsum([len(q) for q in queries])
    qterms = np.array([t for q in queries for t in q], np.uint32) if qoff[-1] else np.zeros(1, np.uint32)
    return qterms, qoff


def eval_queries(ix, queries, nthreads=1, full=False, timings=False):
    """Exact results: (cum_counts[nq], results[...]) [, per-query ms]."""
    qterms, qoff = _queries_flat(queries) if not isinstance(queries, tuple) else queries
    nq = len(qoff) - 1
    cap = 0
    for q in range(nq):
        ts = qterms[int(qoff[q]):int(qoff[q + 1])]
        if len(ts):
            cap += int(min(ix.ucnt[t] for t in ts))
    res = np.zeros(max(cap, 1), np.uint32)
    cum = np.zeros(max(nq, 1), np.uint32)
    qms = np.zeros(max(nq, 1), np.float64) if timings else None
    if full:
        rc = lib().orc_eval_queries_full(ctypes.byref(ix.struct()), _p(qterms), _p(qoff), nq,
                                         _p(cum), _p(res), cap)
    else:
        rc = lib().orc_eval_queries(ctypes.byref(ix.struct()), _p(qterms), _p(qoff), nq, nthreads,
                                    _p(cum), _p(res), cap, _p(qms) if timings else None)
    if rc != 0:
        raise RuntimeError("oracle eval failed")
    n = int(cum[nq - 1]) if nq else 0
    if timings:
        return cum[:nq], res[:n], qms[:nq]
    return cum[:nq], res[:n]


def checksum(cum, res):
    cs = np.zeros(3, np.uint64)
    c = _u32(cum)
    r = _u32(res) if len(res) else np.zeros(1, np.uint32)
    lib().orc_checksum(_p(c), _p(r), len(c), cs.ctypes.data)
    return tuple(int(x) for x in cs)


def query_bytes(ix, queries):
    qterms, qoff = _queries_flat(queries) if not isinstance(queries, tuple) else queries
    full = np.zeros(1, np.uint64)
    b = lib().orc_query_bytes(ctypes.byref(ix.struct()), _p(qterms), _p(qoff), len(qoff) - 1,
                              full.ctypes.data)
    return int(b), int(full[0])


def intersect_all(lists):
    """Brute-force set intersection (numpy), the closed-form judge."""
    if not lists:
        return np.zeros(0, np.uint32)
    acc = _u32(lists[0])
    for x in lists[1:]:
        acc = np.intersect1d(acc, _u32(x), assume_unique=True)
    return acc.astype(np.uint32)


# ------------------------------------------------------------ learned models

def pla_fit(L):
    """(x0, xs, knots[257]) of the PLA-256 position model of sorted list L."""
    a = _u32(L)
    x0 = np.zeros(1, np.float32)
    xs = np.zeros(1, np.float32)
    y = np.zeros(257, np.float32)
    lib().orc_pla_fit(_p(a), len(a), x0.ctypes.data, xs.ctypes.data, y.ctypes.data)
    return float(x0[0]), float(xs[0]), y


def plaq_fit(L):
    """params (x[257], y[257], s[256]) of the PLA-Q position model of sorted
    list L: knots at 256 equal position steps."""
    a = _u32(L)
    prm = np.zeros(770, np.float32)
    lib().orc_plaq_fit(_p(a), len(a), prm.ctypes.data)
    return prm


def plaq_pred(params, v):
    prm = np.ascontiguousarray(params, np.float32)
    return float(lib().orc_plaq_pred(prm.ctypes.data, int(v)))


def model_max_err(kind, params, x0, xs, L):
    a = _u32(L)
    prm = np.ascontiguousarray(params, np.float32)
    return float(lib().orc_model_max_err(kind, prm.ctypes.data, x0, xs, _p(a), len(a)))


def mlp_scale(L):
    x0 = np.float32(L[0])
    xs = np.float32(1.0) / (np.float32(L[-1]) - x0 + np.float32(1.0))
    return float(x0), float(xs)


def mlp_loss_grad(params, L, samples, _plant=None):
    """Mean squared position error of the 1-32-1 ReLU MLP over `samples`
    strided pairs of L and its analytic gradient, in float64.

    _plant (tests only) plants a known slip in the gradient so that a test can
    show its finite-difference check catches it: "relu_mask" drops the ReLU
    derivative, "scale" drops the 2/S factor."""
    L = _u32(L)
    n = len(L)
    S = min(samples, n)
    idx = (np.arange(S, dtype=np.uint64) * np.uint64(n)) // np.uint64(S)
    x0, xs = mlp_scale(L)
    x = ((L[idx].astype(np.float32) - np.float32(x0)) * np.float32(xs)).astype(np.float64)
    t = idx.astype(np.float64) / n
    w = np.asarray(params, np.float64)
    w1, b1, w2, b2 = w[:32], w[32:64], w[64:96], w[96]
    pre = np.outer(x, w1) + b1
    h = np.maximum(pre, 0.0)
    out = h @ w2 + b2
    e = out - t
    loss = float(np.mean(e * e))
    g = e if _plant == "scale" else 2.0 * e / S
    gw2 = h.T @ g
    gb2 = g.sum()
    dh = np.outer(g, w2) * (1.0 if _plant == "relu_mask" else (pre > 0))
    gw1 = dh.T @ x
    gb1 = dh.sum(axis=0)
    return loss, np.concatenate([gw1, gb1, gw2, [gb2]])
