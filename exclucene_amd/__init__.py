"""exclucene_amd -- MI355X-native decode + exact intersection of compressed
posting lists (LRC, SegLRC, HS_SegLRC, ParaPFD), behind the C ABI declared in
include/exclucene_amd.h.

This module is the host-side mirror of the reference's per-program stage
functions (src/gpu/Compressed/src/<codec>/<codec>_Intersection/*.cu:
CPUPreprocess / htodTransfer / kernelInvoke / dtohTransfer / verify) over
ctypes.  All decoding and intersecting happens in the HIP kernels of
libexclucene_amd.so; there is no CPU fallback -- loading fails loudly if the
library is missing.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libexclucene_amd.so")

LRC, SEGLRC, HS_SEGLRC, PARAPFD, NEWPFD, RAW, RAW_HS = 0, 1, 2, 3, 4, 5, 6  # NewPFD: decompression only
CODECS = {"LRC": LRC, "SegLRC": SEGLRC, "HS_SegLRC": HS_SEGLRC, "ParaPFD": PARAPFD, "NewPFD": NEWPFD,
          "uncompressed": RAW, "uncompressed_hash_segment": RAW_HS}

OPT_SEARCH, OPT_TPE, OPT_PERSISTENT, OPT_GRAPH, OPT_TIMING, OPT_COOP_MAX = 0, 1, 2, 3, 4, 5
SEARCH_BINARY, SEARCH_FIT, SEARCH_LEARNED, SEARCH_INTERP = 0, 1, 2, 3

EXPORTS = [
    "exc_ctx_create", "exc_ctx_destroy", "exc_last_error", "exc_ctx_set_option", "exc_ctx_last_times",
    "exc_index_upload", "exc_index_load_files", "exc_index_free", "exc_index_device_bytes", "exc_index_lengths",
    "exc_CPUPreprocess", "exc_htodTransfer", "exc_kernelInvoke", "exc_dtohTransfer", "exc_verify",
    "exc_intersect", "exc_plan", "exc_decompress",
    "exc_build_index", "exc_built_view", "exc_built_free", "exc_index_save_files",
    "exc_synth_lists", "exc_index_train_models", "exc_index_model_info", "exc_mlp_gradient",
    "exc_server_create", "exc_server_submit", "exc_server_wait", "exc_server_replay", "exc_server_stats",
    "exc_server_destroy", "exc_autotune", "exc_ctx_set_tuning", "exc_decompress_device",
    "exc_device_alloc", "exc_device_free", "exc_memcpy",
    "exc_synth_lists_device", "exc_build_index_device", "exc_synth_index_device", "exc_index_export",
    "exc_intersect_noalloc",
    "exc_host_alloc", "exc_host_free",
    "exc_router_create", "exc_router_autotune", "exc_router_set_routes", "exc_router_set_option",
    "exc_router_plan", "exc_router_invoke", "exc_router_fetch", "exc_router_intersect", "exc_router_destroy",
]
MODEL_PLA, MODEL_MLP, MODEL_PLA_Q = 1, 2, 3
SERVER_COUNTS = 1

_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64


class IndexHost(ctypes.Structure):
    """exc_index_host (include/exclucene_amd.h)."""
    _fields_ = [("codec", ctypes.c_int32), ("seg_size", _u32), ("nterms", _u32), ("reserved", _u32),
                ("ind1", _vp), ("ind2", _vp), ("ind2_words", _u64),
                ("seg_num", _vp), ("seg_offset", _vp), ("nseg", _u64),
                ("seg_head", _vp), ("seg_median", _vp), ("lrc_info", _vp), ("seg_info", _vp),
                ("seg_base", _vp), ("list_info", _vp), ("hash_offset", _vp), ("hash_info", _vp),
                ("hash_words", _u64), ("gu_regression", _vp)]


RATIO_CLASSES = 6


class TuneTable(ctypes.Structure):
    _fields_ = [("queries", _u32 * RATIO_CLASSES), ("coop_max", _u32 * RATIO_CLASSES),
                ("search", _u32 * RATIO_CLASSES), ("ms", ctypes.c_float * RATIO_CLASSES),
                ("ms_default", ctypes.c_float * RATIO_CLASSES), ("persistent", _u32),
                ("batch_ms", ctypes.c_float * 2)]

    def as_dict(self):
        return {"queries": list(self.queries), "coop_max": list(self.coop_max), "search": list(self.search),
                "ms": [float(x) for x in self.ms], "ms_default": [float(x) for x in self.ms_default],
                "persistent": int(self.persistent), "batch_ms": [float(x) for x in self.batch_ms]}


ROUTER_MAX_CANDIDATES = 8


class RouteTable(ctypes.Structure):
    _fields_ = [("ncands", _u32), ("route", _u32 * RATIO_CLASSES),
                ("class_ms", (ctypes.c_float * RATIO_CLASSES) * ROUTER_MAX_CANDIDATES),
                ("single_ms", ctypes.c_float * ROUTER_MAX_CANDIDATES), ("routed_ms", ctypes.c_float),
                ("persistent", _u32)]

    def as_dict(self):
        n = int(self.ncands)
        return {"route": list(self.route), "class_ms": [[float(x) for x in self.class_ms[i]] for i in range(n)],
                "single_ms": [float(x) for x in self.single_ms[:n]], "routed_ms": float(self.routed_ms),
                "persistent": int(self.persistent)}


class ServerInfo(ctypes.Structure):
    _fields_ = [("batches", _u64), ("queries", _u64), ("max_inflight_batches", _u64),
                ("overlapped_launches", _u64)]


class ModelStats(ctypes.Structure):
    _fields_ = [("nmodels", _u32), ("max_err", ctypes.c_float), ("mean_err", ctypes.c_float),
                ("train_ms", ctypes.c_float)]


class ModelHost(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("n", _u32), ("err", ctypes.c_float), ("x0", ctypes.c_float),
                ("xs", ctypes.c_float), ("nparams", _u32), ("params", ctypes.c_float * 770)]


class BatchInfo(ctypes.Structure):
    """batchInfo (LRC_Intersection_kernel.cu:68-74)."""
    _fields_ = [("blockNum", _u32), ("constantUsedInByte", _u32), ("ucntShortest_sum", _u32),
                ("nTotalQueryNum", _u32), ("nLinearInfoInByte", _u32)]


class CheckSum(ctypes.Structure):
    """checkSum (LRC_Intersection_kernel.cu:79-83)."""
    _fields_ = [("checkSum1", _u64), ("checkSum2", _u64), ("checkSum3", _u64)]


IND1 = np.dtype([("m_urlcount", "<u4"), ("pad", "<u4"), ("m_off", "<u8")])
LRC_INFO = np.dtype([("fSlope", "<f4"), ("iIntercept", "<i4"), ("fSearchRange", "<f4")])
SEG_INFO = np.dtype([("fSlope", "<f4"), ("fIntercept", "<f4"), ("nRangeLeft", "<u4"), ("nRangeRight", "<u4")])
GU_REG_INFO = np.dtype([("fSlope", "<f4"), ("fIntercept", "<f4"), ("fRSquare", "<f4"), ("fMultiple", "<f4"),
                        ("fInterceptHigh", "<f4"), ("fInterceptLow", "<f4"), ("nRangeLeft", "<u4"),
                        ("nRangeRight", "<u4"), ("nMagnification", "<u4")])

_lib = None


class ExcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("exclucene_amd error %d: %s" % (code, msg))
        self.code = code


def lib():
    """Load libexclucene_amd.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libexclucene_amd.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name in EXPORTS:
            getattr(L, name).restype = ctypes.c_int
        L.exc_last_error.restype = ctypes.c_char_p
        L.exc_ctx_destroy.restype = None
        L.exc_index_free.restype = None
        L.exc_built_free.restype = None
        L.exc_index_device_bytes.restype = _u64
        L.exc_ctx_create.argtypes = [ctypes.c_int, _u64, ctypes.POINTER(_vp)]
        L.exc_ctx_destroy.argtypes = [_vp]
        L.exc_ctx_set_option.argtypes = [_vp, ctypes.c_int, ctypes.c_int64]
        L.exc_ctx_last_times.argtypes = [_vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
        L.exc_index_upload.argtypes = [_vp, ctypes.POINTER(IndexHost), ctypes.POINTER(_vp)]
        L.exc_index_load_files.argtypes = [_vp, ctypes.c_int, _u32, ctypes.c_char_p, ctypes.c_char_p,
                                           ctypes.POINTER(_vp)]
        L.exc_index_free.argtypes = [_vp]
        L.exc_index_device_bytes.argtypes = [_vp]
        L.exc_index_lengths.argtypes = [_vp, _vp, _u32, ctypes.POINTER(_u32)]
        L.exc_CPUPreprocess.argtypes = [_vp, _vp, ctypes.POINTER(_vp), _vp, _u64, ctypes.POINTER(BatchInfo)]
        L.exc_htodTransfer.argtypes = [_vp, ctypes.POINTER(BatchInfo)]
        L.exc_kernelInvoke.argtypes = [_vp, ctypes.POINTER(BatchInfo)]
        L.exc_dtohTransfer.argtypes = [_vp, ctypes.POINTER(BatchInfo), _vp, _vp, _u64]
        L.exc_verify.argtypes = [ctypes.POINTER(BatchInfo), _vp, _vp, ctypes.POINTER(CheckSum)]
        L.exc_intersect.argtypes = [_vp, _vp, _vp, _vp, _u32, _vp, _vp, _u64, ctypes.POINTER(_u64)]
        L.exc_intersect_noalloc.argtypes = [_vp, _vp, _vp, _vp, _u32, _vp, _vp, _u64, ctypes.POINTER(_u64)]
        L.exc_plan.argtypes = [_vp, _vp, _vp, _vp, _u32, ctypes.POINTER(BatchInfo)]
        L.exc_decompress.argtypes = [_vp, _vp, _vp, _u32, _vp, _u64, ctypes.POINTER(_u64)]
        L.exc_build_index.argtypes = [ctypes.c_int, _u32, ctypes.c_float, _u32, _u32, _vp, _vp, _u32,
                                      ctypes.c_int, ctypes.POINTER(_vp)]
        L.exc_built_view.argtypes = [_vp, ctypes.POINTER(IndexHost)]
        L.exc_built_free.argtypes = [_vp]
        L.exc_index_save_files.argtypes = [ctypes.POINTER(IndexHost), ctypes.c_char_p, ctypes.c_char_p]
        L.exc_synth_lists.argtypes = [_u64, _u32, _vp, _u32, ctypes.c_int, _vp, _vp]
        L.exc_index_train_models.argtypes = [_vp, _vp, ctypes.c_int, _u32, _u32, _u32, ctypes.c_float, _u64,
                                             ctypes.POINTER(ModelStats)]
        L.exc_index_model_info.argtypes = [_vp, _vp, _u32, ctypes.POINTER(ModelHost)]
        L.exc_mlp_gradient.argtypes = [_vp, _vp, _vp, _u32, _u32, _vp, ctypes.POINTER(ctypes.c_float)]
        L.exc_synth_lists_device.argtypes = [_vp, _u64, _u32, _vp, _u32, _vp, _vp, _u64]
        L.exc_build_index_device.argtypes = [_vp, ctypes.c_int, _u32, ctypes.c_float, _u32, _u32, _vp, _vp, _u32,
                                             ctypes.POINTER(_vp)]
        L.exc_synth_index_device.argtypes = [_vp, ctypes.c_int, _u32, ctypes.c_float, _u32, _u32, _u64, _u32, _vp,
                                             _u32, _u64, _vp, ctypes.POINTER(_vp)]
        L.exc_index_export.argtypes = [_vp, _vp, ctypes.POINTER(_vp)]
        L.exc_device_alloc.argtypes = [_vp, _u64, ctypes.POINTER(_vp)]
        L.exc_device_free.argtypes = [_vp, _vp]
        L.exc_memcpy.argtypes = [_vp, _vp, _vp, _u64, ctypes.c_int]
        L.exc_host_alloc.argtypes = [_vp, _u64, ctypes.POINTER(_vp)]
        L.exc_host_free.argtypes = [_vp, _vp]
        L.exc_decompress_device.argtypes = [_vp, _vp, _vp, _u32, _vp, _u64, ctypes.POINTER(_u64)]
        L.exc_autotune.argtypes = [_vp, _vp, _vp, _vp, _u32, _u32, ctypes.POINTER(TuneTable)]
        L.exc_ctx_set_tuning.argtypes = [_vp, ctypes.POINTER(TuneTable)]
        L.exc_router_create.argtypes = [_vp, _vp, _u32, ctypes.POINTER(_vp)]
        L.exc_router_autotune.argtypes = [_vp, _vp, _vp, _u32, _u32, ctypes.POINTER(RouteTable)]
        L.exc_router_set_routes.argtypes = [_vp, _vp]
        L.exc_router_set_option.argtypes = [_vp, ctypes.c_int, ctypes.c_int64]
        L.exc_router_plan.argtypes = [_vp, _vp, _vp, _u32]
        L.exc_router_invoke.argtypes = [_vp, ctypes.POINTER(ctypes.c_float)]
        L.exc_router_fetch.argtypes = [_vp, _vp, _vp, _u64, ctypes.POINTER(_u64)]
        L.exc_router_intersect.argtypes = [_vp, _vp, _vp, _u32, _vp, _vp, _u64, ctypes.POINTER(_u64)]
        L.exc_router_destroy.argtypes = [_vp]
        L.exc_router_destroy.restype = None
        L.exc_server_destroy.restype = None
        L.exc_server_create.argtypes = [_vp, _vp, _u32, _u32, _u64, _u32, ctypes.POINTER(_vp)]
        L.exc_server_submit.argtypes = [_vp, _vp, _u32, ctypes.POINTER(_u64)]
        L.exc_server_wait.argtypes = [_vp, _u64, _vp, _u64, ctypes.POINTER(_u64), ctypes.POINTER(ctypes.c_double)]
        L.exc_server_replay.argtypes = [_vp, _vp, _vp, _vp, _u32, _vp, _vp, _vp, _u64]
        L.exc_server_stats.argtypes = [_vp, ctypes.POINTER(ServerInfo)]
        L.exc_server_destroy.argtypes = [_vp]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise ExcError(rc, lib().exc_last_error().decode(errors="replace"))


def _p(a):
    return a.ctypes.data if a is not None and a.size else None


def _arr(ptr, count, dtype):
    if not ptr or count == 0:
        return None
    dt = np.dtype(dtype)
    buf = (ctypes.c_char * (count * dt.itemsize)).from_address(ptr)
    return np.frombuffer(buf, dtype=dt, count=count)


# --------------------------------------------------------------- corpus + build

def synth_lists(df, universe, seed, nthreads=0):
    """Sorted uniform samples without replacement, list t of length df[t]."""
    df = np.ascontiguousarray(df, np.uint32)
    off = np.zeros(len(df) + 1, np.uint64)
    off[1:] = np.cumsum(df, dtype=np.uint64)
    docs = np.empty(max(int(off[-1]), 1), np.uint32)
    _check(lib().exc_synth_lists(seed, universe, _p(df), len(df), nthreads, _p(docs), _p(off)))
    return docs[:int(off[-1])], off


class BuiltIndex:
    """A compressed index built on the host (the *_Generator + *_Compression
    programs' output), viewed as numpy arrays without copies."""

    def __init__(self, handle):
        self._h = handle
        self.view = IndexHost()
        _check(lib().exc_built_view(handle, ctypes.byref(self.view)))
        v = self.view
        nt, ns = v.nterms, v.nseg
        self.codec, self.seg, self.nterms, self.nseg = v.codec, v.seg_size, nt, ns
        self.ind1 = _arr(v.ind1, nt, IND1)
        self.ind2 = _arr(v.ind2, v.ind2_words, np.uint32)
        self.seg_num = _arr(v.seg_num, nt, np.uint32)
        self.seg_offset = _arr(v.seg_offset, ns, np.uint32)
        self.seg_head = _arr(v.seg_head, ns, np.uint32)
        self.seg_median = _arr(v.seg_median, ns, np.uint32)
        self.lrc_info = _arr(v.lrc_info, nt, LRC_INFO)
        self.seg_info = _arr(v.seg_info, ns, SEG_INFO)
        self.seg_base = _arr(v.seg_base, ns, np.int32)
        self.list_info = _arr(v.list_info, nt, SEG_INFO)
        self.hash_offset = _arr(v.hash_offset, nt, np.uint32)
        self.hash_info = _arr(v.hash_info, v.hash_words, np.uint32)
        self.gu_regression = _arr(v.gu_regression, nt, GU_REG_INFO)

    @property
    def ucnt(self):
        return self.ind1["m_urlcount"]

    def compressed_bytes(self):
        tot = 0
        for name in ("ind1", "ind2", "seg_num", "seg_offset", "seg_head", "seg_median", "lrc_info",
                     "seg_info", "seg_base", "hash_offset", "hash_info"):
            a = getattr(self, name)
            if a is not None:
                tot += a.nbytes
        return tot

    def save(self, directory, name):
        _check(lib().exc_index_save_files(ctypes.byref(self.view), directory.encode(), name.encode()))

    def close(self):
        if self._h:
            lib().exc_built_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def build_index(docs, offsets, codec, seg=128, frac=0.0, hs_target=256, hs_shift=0, nthreads=0):
    """Encode concatenated sorted lists (docs, offsets[nterms+1]) into one index."""
    docs = np.ascontiguousarray(docs, np.uint32)
    offsets = np.ascontiguousarray(offsets, np.uint64)
    h = _vp()
    _check(lib().exc_build_index(codec, seg, frac, hs_target, hs_shift, _p(docs) or docs.ctypes.data,
                                 _p(offsets), len(offsets) - 1, nthreads, ctypes.byref(h)))
    return BuiltIndex(h)


def lists_to_flat(lists):
    lists = [np.ascontiguousarray(x, np.uint32) for x in lists]
    off = np.zeros(len(lists) + 1, np.uint64)
    off[1:] = np.cumsum([len(x) for x in lists])
    docs = np.concatenate(lists) if off[-1] else np.zeros(1, np.uint32)
    return docs, off


def queries_to_flat(queries):
    qoff = np.zeros(len(queries) + 1, np.uint64)
    qoff[1:] = np.cumsum([len(q) for q in queries])
    qterms = np.array([t for q in queries for t in q], np.uint32) if qoff[-1] else np.zeros(1, np.uint32)
    return qterms, qoff


def queries_to_stream(queries):
    """The reference's .query byte stream: {u32 tnum; u32 tno[tnum]}..."""
    out = []
    for q in queries:
        out.append(len(q))
        out.extend(q)
    return np.array(out, np.uint32).tobytes()


# --------------------------------------------------------------- device side

class DeviceIndex:
    def __init__(self, ctx, handle, built=None, lengths=None):
        self.ctx, self._h, self.built = ctx, handle, built
        self._lengths = lengths

    @property
    def ucnt(self):
        """List lengths, or None when unknown (an index loaded from files)."""
        if self.built is not None:
            return self.built.ucnt
        return self._lengths

    @property
    def handle(self):
        return self._h

    def device_bytes(self):
        return int(lib().exc_index_device_bytes(self._h))

    def close(self):
        if self._h:
            lib().exc_index_free(self._h)
            self._h = None


def _result_cap(ucnt, qterms, qoff):
    """Results a batch can hold at most: the summed shortest-list lengths
    (terms out of range count as empty lists; the library reports them)."""
    if ucnt is None:
        return 1 << 26
    qoff = np.asarray(qoff, np.int64)
    nq = len(qoff) - 1
    if nq <= 0 or qoff[-1] == 0:
        return 0
    t = np.asarray(qterms, np.int64)[:qoff[-1]]
    n = np.where(t < len(ucnt), ucnt[np.minimum(t, len(ucnt) - 1)].astype(np.int64), 0)
    starts = qoff[:-1]
    nonempty = qoff[1:] > starts
    mins = np.minimum.reduceat(n, np.minimum(starts, len(n) - 1))
    return int(mins[nonempty].sum())


class _Pinned:
    """One exc_host_alloc buffer, freed when the last array over it goes."""

    def __init__(self, ctx, nbytes):
        p = _vp()
        _check(lib().exc_host_alloc(ctx.handle, nbytes, ctypes.byref(p)))
        self.ctx, self.ptr, self.nbytes = ctx, p.value, nbytes

    def __del__(self):
        try:
            if self.ptr and self.ctx._h:
                lib().exc_host_free(self.ctx._h, self.ptr)
        except Exception:
            pass
        self.ptr = None


class Context:
    """One HIP stream and its batch buffers (one per host thread)."""

    def __init__(self, device=0, budget=0):
        h = _vp()
        _check(lib().exc_ctx_create(device, budget, ctypes.byref(h)))
        self._h = h
        self._bi = None

    @property
    def handle(self):
        return self._h

    def set_option(self, key, value):
        _check(lib().exc_ctx_set_option(self._h, key, int(value)))

    def last_times(self):
        a, b = ctypes.c_float(), ctypes.c_float()
        _check(lib().exc_ctx_last_times(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def upload(self, built):
        h = _vp()
        _check(lib().exc_index_upload(self._h, ctypes.byref(built.view), ctypes.byref(h)))
        return DeviceIndex(self, h, built)

    def load_files(self, codec, seg, directory, name):
        h = _vp()
        _check(lib().exc_index_load_files(self._h, codec, seg, directory.encode(), name.encode(), ctypes.byref(h)))
        n = _u32()
        _check(lib().exc_index_lengths(h, None, 0, ctypes.byref(n)))
        lengths = np.zeros(max(int(n.value), 1), np.uint32)
        _check(lib().exc_index_lengths(h, _p(lengths), len(lengths), ctypes.byref(n)))
        return DeviceIndex(self, h, lengths=lengths[:int(n.value)])

    # -- reference stage API
    def CPUPreprocess(self, index, stream_bytes, pos, threshold):
        """Plan one batch from the .query stream starting at byte pos; returns
        (batchInfo, new_pos)."""
        buf = ctypes.create_string_buffer(stream_bytes, len(stream_bytes)) if isinstance(stream_bytes, bytes) else stream_bytes
        base = ctypes.addressof(buf)
        p = _vp(base + pos)
        bi = BatchInfo()
        _check(lib().exc_CPUPreprocess(self._h, index.handle, ctypes.byref(p), base + len(stream_bytes),
                                       threshold, ctypes.byref(bi)))
        self._bi = bi
        return bi, p.value - base

    def htodTransfer(self, bi):
        _check(lib().exc_htodTransfer(self._h, ctypes.byref(bi)))

    def kernelInvoke(self, bi):
        _check(lib().exc_kernelInvoke(self._h, ctypes.byref(bi)))

    def dtohTransfer(self, bi, cap=None, out=None):
        """(cum_counts, results) of the invoked batch; out = (cum, res)
        buffers to fill (page-locked ones from host_empty copy fastest),
        else new arrays."""
        nq = bi.nTotalQueryNum
        cap = bi.ucntShortest_sum if cap is None else cap
        cum, res = self._out_arrays(nq, cap, out)
        cap = min(cap, len(res))
        _check(lib().exc_dtohTransfer(self._h, ctypes.byref(bi), _p(cum), _p(res), cap))
        n = int(cum[nq - 1]) if nq else 0
        return cum[:nq], res[:n]

    @staticmethod
    def verify(bi, cum, res):
        cs = CheckSum()
        c = np.ascontiguousarray(cum, np.uint32)
        r = np.ascontiguousarray(res, np.uint32) if len(res) else np.zeros(1, np.uint32)
        _check(lib().exc_verify(ctypes.byref(bi), _p(c), _p(r), ctypes.byref(cs)))
        return cs.checkSum1, cs.checkSum2, cs.checkSum3

    # -- direct API
    def plan(self, index, queries):
        qterms, qoff = queries if isinstance(queries, tuple) else queries_to_flat(queries)
        bi = BatchInfo()
        _check(lib().exc_plan(self._h, index.handle, _p(qterms), _p(qoff), len(qoff) - 1, ctypes.byref(bi)))
        self._bi = bi
        return bi

    def intersect(self, index, queries, out=None):
        """Exact results of every query: (cum_counts, results), filled into
        out = (cum, res) when given (views of them are returned)."""
        return self._intersect_with(lib().exc_intersect, index, queries, out)

    def intersect_noalloc(self, index, queries):
        """As intersect, through exc_intersect_noalloc: no allocation, the
        batch planned on the device inside the library's fixed scratch area."""
        return self._intersect_with(lib().exc_intersect_noalloc, index, queries)

    @staticmethod
    def _out_arrays(nq, cap, out):
        if out is None:
            return np.zeros(max(nq, 1), np.uint32), np.zeros(max(cap, 1), np.uint32)
        cum, res = out
        if len(cum) < max(nq, 1) or cum.dtype != np.uint32 or res.dtype != np.uint32 or len(res) < 1:
            raise ValueError("out buffers: uint32, at least one count per query")
        return cum, res

    def host_empty(self, count, dtype=np.uint32):
        """A numpy array of `count` elements in page-locked host memory
        (exc_host_alloc), freed when the array is garbage (or with the
        context)."""
        dt = np.dtype(dtype)
        pin = _Pinned(self, max(1, int(count)) * dt.itemsize)
        buf = (ctypes.c_char * pin.nbytes).from_address(pin.ptr)
        buf._pin = pin  # the array's base keeps the allocation alive
        return np.frombuffer(buf, dtype=dt, count=int(count))

    def _intersect_with(self, fn, index, queries, out=None):
        qterms, qoff = queries if isinstance(queries, tuple) else queries_to_flat(queries)
        nq = len(qoff) - 1
        cap = _result_cap(index.ucnt, qterms, qoff)
        cum, res = self._out_arrays(nq, cap, out)
        cap = min(cap, len(res))
        n = _u64()
        _check(fn(self._h, index.handle, _p(qterms), _p(qoff), nq, _p(cum), _p(res), cap, ctypes.byref(n)))
        return cum[:nq], res[:int(n.value)]

    # -- autotuner
    def autotune(self, index, queries, reps=3):
        """Pick per-ratio-class knobs and the launch mode on this sample and
        install them (exc_autotune); returns the table as a dict."""
        qterms, qoff = queries if isinstance(queries, tuple) else queries_to_flat(queries)
        t = TuneTable()
        _check(lib().exc_autotune(self._h, index.handle, _p(qterms), _p(qoff), len(qoff) - 1, reps,
                                  ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, table):
        """Install a tuning table (dict with coop_max, search lists and persistent), or None."""
        if table is None:
            _check(lib().exc_ctx_set_tuning(self._h, None))
            return
        t = TuneTable()
        for k in range(RATIO_CLASSES):
            t.coop_max[k] = int(table["coop_max"][k])
            t.search[k] = int(table["search"][k])
        t.persistent = int(table.get("persistent", 0))
        _check(lib().exc_ctx_set_tuning(self._h, ctypes.byref(t)))

    # -- learned position models
    def train_models(self, index, kind=MODEL_PLA, min_len=65536, epochs=300, samples=4096, lr=3e-4, seed=4):
        st = ModelStats()
        _check(lib().exc_index_train_models(self._h, index.handle, kind, min_len, epochs, samples, lr, seed,
                                            ctypes.byref(st)))
        return {"nmodels": st.nmodels, "max_err": st.max_err, "mean_err": st.mean_err, "train_ms": st.train_ms}

    def model_info(self, index, term):
        m = ModelHost()
        _check(lib().exc_index_model_info(self._h, index.handle, term, ctypes.byref(m)))
        return {"kind": m.kind, "n": m.n, "err": m.err, "x0": m.x0, "xs": m.xs,
                "params": np.array(m.params[:m.nparams], np.float32)}

    def mlp_gradient(self, params, docs, samples):
        p = np.ascontiguousarray(params, np.float32)
        d = np.ascontiguousarray(docs, np.uint32)
        g = np.zeros(97, np.float32)
        loss = ctypes.c_float()
        _check(lib().exc_mlp_gradient(self._h, _p(p), _p(d), len(d), samples, _p(g), ctypes.byref(loss)))
        return g, loss.value

    def decompress(self, index, terms):
        terms = np.ascontiguousarray(terms, np.uint32)
        ucnt = index.ucnt
        cap = int(sum(int(ucnt[t]) for t in terms))
        out = np.zeros(max(cap, 1), np.uint32)
        n = _u64()
        _check(lib().exc_decompress(self._h, index.handle, _p(terms), len(terms), _p(out), cap, ctypes.byref(n)))
        return out[:int(n.value)]

    # -- index construction on the GPU
    def synth_lists_device(self, seed, universe, df, d_out=None, cap=0):
        """Seeded Bernoulli(df/universe) lists on the device; returns the offsets
        (and writes the docIDs to device address d_out when given)."""
        df = np.ascontiguousarray(df, np.uint32)
        off = np.zeros(len(df) + 1, np.uint64)
        _check(lib().exc_synth_lists_device(self._h, seed, universe, _p(df), len(df), _p(off),
                                            ctypes.c_void_p(d_out) if d_out else None, cap))
        return off

    def build_index_device(self, d_docs, list_off, seg=128, codec=SEGLRC, frac=0.0, hs_target=256, hs_shift=0):
        """Encode sorted lists already in device memory (SegLRC, LRC, ParaPFD or HS_SegLRC)."""
        off = np.ascontiguousarray(list_off, np.uint64)
        h = _vp()
        _check(lib().exc_build_index_device(self._h, codec, seg, frac, hs_target, hs_shift, ctypes.c_void_p(d_docs),
                                            _p(off), len(off) - 1, ctypes.byref(h)))
        return DeviceIndex(self, h, lengths=np.diff(off).astype(np.uint32))

    def synth_index_device(self, seed, universe, df, seg=128, batch_postings=1 << 30, codec=SEGLRC, frac=0.0,
                           hs_target=256, hs_shift=0):
        """Synthesize and encode an index (SegLRC, LRC, ParaPFD or HS_SegLRC)
        on the device in batches; returns (DeviceIndex, list offsets)."""
        df = np.ascontiguousarray(df, np.uint32)
        off = np.zeros(len(df) + 1, np.uint64)
        h = _vp()
        _check(lib().exc_synth_index_device(self._h, codec, seg, frac, hs_target, hs_shift, seed, universe, _p(df),
                                            len(df), batch_postings, _p(off), ctypes.byref(h)))
        return DeviceIndex(self, h, lengths=np.diff(off).astype(np.uint32)), off

    def export(self, index):
        """The index's arrays in the reference layout on the host (SegLRC, LRC, ParaPFD)."""
        h = _vp()
        _check(lib().exc_index_export(self._h, index.handle, ctypes.byref(h)))
        return BuiltIndex(h)

    # -- device buffers owned by the library
    def device_alloc(self, nbytes):
        p = _vp()
        _check(lib().exc_device_alloc(self._h, nbytes, ctypes.byref(p)))
        return int(p.value)

    def device_free(self, d_ptr):
        _check(lib().exc_device_free(self._h, ctypes.c_void_p(d_ptr)))

    def to_device(self, d_ptr, arr):
        a = np.ascontiguousarray(arr)
        _check(lib().exc_memcpy(self._h, ctypes.c_void_p(d_ptr), _p(a), a.nbytes, 1))

    def from_device(self, d_ptr, count, dtype=np.uint32):
        out = np.empty(count, dtype)
        _check(lib().exc_memcpy(self._h, _p(out), ctypes.c_void_p(d_ptr), out.nbytes, 0))
        return out

    def decompress_device(self, index, terms, d_out, cap):
        """Decode whole lists into device memory at address d_out (capacity
        `cap` docIDs); returns the docIDs written."""
        terms = np.ascontiguousarray(terms, np.uint32)
        n = _u64()
        _check(lib().exc_decompress_device(self._h, index.handle, _p(terms), len(terms), ctypes.c_void_p(d_out), cap,
                                           ctypes.byref(n)))
        return int(n.value)

    def close(self):
        if self._h:
            lib().exc_ctx_destroy(self._h)
            self._h = None


class Router:
    """One batch routed by ratio class over several indexes of the same lists
    (exc_router_*): each candidate's sub-batch on a stream of its own, the
    whole run timed as one."""

    def __init__(self, ctx, indexes):
        self.ctx = ctx
        self.indexes = list(indexes)
        arr = (_vp * len(self.indexes))(*[ix.handle.value for ix in self.indexes])
        h = _vp()
        _check(lib().exc_router_create(ctx.handle, arr, len(self.indexes), ctypes.byref(h)))
        self._h = h
        self._cap = 0
        self._nq = 0
        ucnts = [ix.ucnt for ix in self.indexes if ix.ucnt is not None]
        self._ucnt = ucnts[0] if ucnts else None

    def autotune(self, queries, reps=3):
        qterms, qoff = queries if isinstance(queries, tuple) else queries_to_flat(queries)
        t = RouteTable()
        _check(lib().exc_router_autotune(self._h, _p(qterms), _p(qoff), len(qoff) - 1, reps, ctypes.byref(t)))
        return t.as_dict()

    def set_routes(self, routes):
        r = np.ascontiguousarray(routes, np.uint32)
        assert len(r) == RATIO_CLASSES
        _check(lib().exc_router_set_routes(self._h, _p(r)))

    def set_option(self, key, value):
        _check(lib().exc_router_set_option(self._h, key, int(value)))

    def plan(self, queries):
        qterms, qoff = queries if isinstance(queries, tuple) else queries_to_flat(queries)
        _check(lib().exc_router_plan(self._h, _p(qterms), _p(qoff), len(qoff) - 1))
        self._nq = len(qoff) - 1
        self._cap = _result_cap(self._ucnt, qterms, qoff)

    def invoke(self):
        """Run the planned batch; returns the run's device time in ms."""
        ms = ctypes.c_float()
        _check(lib().exc_router_invoke(self._h, ctypes.byref(ms)))
        return ms.value

    def fetch(self):
        cum = np.zeros(max(self._nq, 1), np.uint32)
        res = np.zeros(max(self._cap, 1), np.uint32)
        n = _u64()
        _check(lib().exc_router_fetch(self._h, _p(cum), _p(res), self._cap, ctypes.byref(n)))
        return cum[:self._nq], res[:int(n.value)]

    def intersect(self, queries):
        """Exact results of every query, in query order: (cum_counts, results)."""
        self.plan(queries)
        self.invoke()
        return self.fetch()

    def close(self):
        if self._h:
            lib().exc_router_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request loop over a resident index (exc_server_*).

    Queries are admitted into `nslots` batch slots, each on its own HIP stream,
    while the device memory of the in-flight batches stays within `mem_budget`
    bytes (0 = unlimited).  The slots take the search options of `ctx`."""

    def __init__(self, ctx, index, nslots=2, max_batch=1024, mem_budget=0, counts_only=False):
        self.index = index
        self.counts_only = counts_only
        h = _vp()
        _check(lib().exc_server_create(ctx.handle, index.handle, nslots, max_batch, mem_budget,
                                       SERVER_COUNTS if counts_only else 0, ctypes.byref(h)))
        self._h = h

    def submit(self, terms):
        t = np.ascontiguousarray(terms, np.uint32)
        tk = _u64()
        _check(lib().exc_server_submit(self._h, _p(t), len(t), ctypes.byref(tk)))
        return int(tk.value)

    def wait(self, ticket, cap=None):
        if cap is None:
            cap = int(self.index.ucnt.max()) if self.index.ucnt is not None else 1 << 26
        out = np.zeros(max(cap, 1), np.uint32)
        n, lat = _u64(), ctypes.c_double()
        _check(lib().exc_server_wait(self._h, ticket, _p(out), cap, ctypes.byref(n), ctypes.byref(lat)))
        count = int(n.value)
        return (out[:0] if self.counts_only else out[:count]), lat.value, count

    def replay(self, queries, arrival_s, cap):
        """Submit every query at its arrival time (seconds from now) and wait
        for all; returns (latency_ms, counts, concatenated results)."""
        qterms, qoff = queries if isinstance(queries, tuple) else queries_to_flat(queries)
        nq = len(qoff) - 1
        arr = np.ascontiguousarray(arrival_s, np.float64)
        assert len(arr) == nq
        lat = np.zeros(max(nq, 1), np.float64)
        cnt = np.zeros(max(nq, 1), np.uint32)
        res = np.zeros(max(cap, 1), np.uint32) if cap else None  # cap 0: counts and latencies only
        _check(lib().exc_server_replay(self._h, _p(qterms), _p(qoff), _p(arr), nq, _p(lat), _p(cnt),
                                       _p(res) if cap else None, cap))
        return lat[:nq], cnt[:nq], (res[:int(cnt[:nq].sum())] if cap else np.zeros(0, np.uint32))

    def stats(self):
        st = ServerInfo()
        _check(lib().exc_server_stats(self._h, ctypes.byref(st)))
        return {k: int(getattr(st, k)) for k, _ in ServerInfo._fields_}

    def close(self):
        if self._h:
            lib().exc_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
