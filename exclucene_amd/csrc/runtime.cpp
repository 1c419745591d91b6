// runtime.cpp -- host runtime behind the C ABI: contexts, device-resident
// indexes, batch planning (CPUPreprocess), launches, result transfer.
//
// Replaces the reference's per-program globals and host stages
// (GC = src/gpu/Compressed/src):
//   allocateResource / releaseResource  GC/LRC/LRC_Intersection/LRC_Intersection.cu:77-122,382-446
//   CPUPreprocess                       :126-246
//   htodTransfer / kernelInvoke         :249-345
//   dtohTransfer / verify               :348-379
// Device memory: the compressed index is uploaded once and stays resident
// (index arrays + one 64-byte record per term); per-batch buffers grow
// geometrically and are reused, so a steady stream of batches allocates nothing.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdarg>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include "exc_format.h"
#include "exc_internal.h"
#include "exc_model.h"
#include "exc_model.h"

using namespace exc;


#define HIP_TRY(call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return set_error(EXC_ERR_DEVICE, "%s failed: %s", #call, hipGetErrorString(e_));   \
    } while (0)

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

}  // namespace

struct exc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t budget = 0;
    uint64_t used = 0;
    int cus = 0;
    // options
    int search = SEARCH_FIT;
    int tpe = 0;
    int persistent = 1;
    int graph = 0;
    int timing = 1;
    int coop_max = 64;
    // batch plan (host)
    exc_index* ix = nullptr;
    std::vector<QueryRec> queries;
    std::vector<uint32_t> qterms;  // sorted terms of the batch
    std::vector<TermRec> qrecs;    // their TermRec copies
    std::vector<UnitRec> units;   // query order: the scan's
    std::vector<UnitRec> launch;  // launch order: the prologue's and the probe's
    uint64_t short_sum = 0;
    uint32_t nq = 0, nunits = 0, seg = 0;
    int state = 0;  // 0 none, 1 planned, 2 on device, 3 invoked
    // device batch buffers
    DevBuf d_queries, d_qrecs, d_units, d_stage, d_ucount, d_uoff, d_bsum, d_result, d_qcum, d_counter;
    DevBuf d_launch;  // the units again, in launch order (build_launch)
    // timing
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    // around the probe inside a step: timing only, without the system-scope
    // fence (its L2 writeback and invalidation otherwise stall the kernels
    // around it); the step's own events ev[0] / ev[3] keep the fence
    hipEvent_t pev[2] = {nullptr, nullptr};
    void* ticket_zeroed = nullptr;  // d_counter allocation whose scan ticket was zeroed
    float probe_ms = 0, invoke_ms = 0;
    // graph of the current plan
    hipGraphExec_t gexec = nullptr;
    // per-ratio-class knobs chosen by exc_autotune (used when tuned)
    bool tuned = false;
    exc_tune_table tune{};
    uint64_t opt_gen = 0, plan_gen = 0;  // knob options changed since the plan was made
    uint64_t plan_model_gen = 0;         // model generation of the batch's TermRec copies
    // device buffers handed to the caller (exc_device_alloc)
    std::unordered_map<void*, size_t> user_bufs;
    // page-locked host buffers handed to the caller (exc_host_alloc)
    std::unordered_map<void*, size_t> host_bufs;
};

struct exc_index {
    exc_ctx* ctx = nullptr;
    int codec = 0;
    uint32_t seg = 256;
    uint32_t nterms = 0;
    uint64_t nseg = 0;
    uint64_t nwords = 0;           // compressed words (.ind2)
    uint64_t hash_words = 0;       // .HashSeg_info words (HS codecs)
    std::vector<uint32_t> n;       // list lengths (host copy for planning)
    std::vector<TermRec> recs;     // host copy of the device term records
    std::vector<uint32_t> lwords;  // .ind2 words of each list (empty: unknown)
    TermRec* d_terms = nullptr;
    std::vector<Model> models;     // learned models (host copy)
    DevBuf d_models, d_params;
    uint64_t model_gen = 0;        // bumped whenever the learned models change
    std::vector<DevBuf> bufs;
    uint64_t bytes = 0;
    DevIndex dix{};
};

namespace {

int dev_alloc(exc_ctx* c, DevBuf& b, size_t bytes)
{
    if (b.bytes >= bytes && b.p) return EXC_OK;
    size_t want = std::max(bytes, b.bytes + b.bytes / 2);
    if (want < 256) want = 256;
    if (c->budget && c->used - b.bytes + want > c->budget) {
        want = bytes < 256 ? 256 : bytes;
        if (c->used - b.bytes + want > c->budget)
            return set_error(EXC_ERR_BUDGET, "device-memory budget of %llu bytes exceeded",
                             (unsigned long long)c->budget);
    }
    if (b.p) {
        (void)hipFree(b.p);
        c->used -= b.bytes;
        b.p = nullptr;
        b.bytes = 0;
    }
    HIP_TRY(hipMalloc(&b.p, want));
    b.bytes = want;
    c->used += want;
    return EXC_OK;
}

void dev_free(exc_ctx* c, DevBuf& b)
{
    if (b.p) {
        (void)hipFree(b.p);
        c->used -= b.bytes;
    }
    b.p = nullptr;
    b.bytes = 0;
}

void drop_graph(exc_ctx* c)
{
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
}

template <class T>
T* dptr(DevBuf& b) { return static_cast<T*>(b.p); }

// upload `bytes` of src followed by kPadWords zero words
// Host-to-device copy of a large host array (a mapped index file, for one)
// through pinned staging.  kStreamThreads host threads each own two staging
// buffers; thread t fills chunks t, t + T, ... from the source and queues each
// chunk's DMA on the context stream, so reading the next chunks (page faults
// of a mapped file included) overlaps the transfers of the previous ones.
constexpr size_t kStreamChunk = size_t(8) << 20;
constexpr int kStreamThreads = 4;

int stream_to_device(exc_ctx* c, void* dst, const void* src, size_t bytes)
{
    const size_t nchunk = (bytes + kStreamChunk - 1) / kStreamChunk;
    const int T = (int)std::min<size_t>(kStreamThreads, nchunk);
    std::vector<void*> stage(2 * T, nullptr);
    std::vector<hipEvent_t> ev(2 * T, nullptr);
    hipError_t err = hipSuccess;
    for (size_t i = 0; i < stage.size() && err == hipSuccess; ++i) {
        err = hipHostMalloc(&stage[i], kStreamChunk, hipHostMallocDefault);
        if (err == hipSuccess) err = hipEventCreateWithFlags(&ev[i], hipEventDisableTiming);
    }
    if (err == hipSuccess) {
        std::vector<hipError_t> terr(T, hipSuccess);
        std::atomic<bool> stop{false};
        std::vector<std::thread> th;
        for (int t = 0; t < T; ++t)
            th.emplace_back([&, t] {
                hipError_t e = hipSetDevice(c->device);
                for (size_t k = t, j = 0; e == hipSuccess && k < nchunk && !stop; k += T, ++j) {
                    const int b = 2 * t + (int)(j & 1);
                    if (j >= 2 && (e = hipEventSynchronize(ev[b])) != hipSuccess) break;  // its last DMA is done
                    const size_t off = k * kStreamChunk, n = std::min(kStreamChunk, bytes - off);
                    std::memcpy(stage[b], static_cast<const char*>(src) + off, n);
                    e = hipMemcpyAsync(static_cast<char*>(dst) + off, stage[b], n, hipMemcpyHostToDevice, c->stream);
                    if (e == hipSuccess) e = hipEventRecord(ev[b], c->stream);
                }
                if (e != hipSuccess) stop = true;
                terr[t] = e;
            });
        for (auto& x : th) x.join();
        err = hipStreamSynchronize(c->stream);
        for (hipError_t e : terr)
            if (e != hipSuccess) err = e;
    }
    for (size_t i = 0; i < stage.size(); ++i) {
        if (ev[i]) (void)hipEventDestroy(ev[i]);
        if (stage[i]) (void)hipHostFree(stage[i]);
    }
    if (err != hipSuccess) return set_error(EXC_ERR_DEVICE, "staged upload: %s", hipGetErrorString(err));
    return EXC_OK;
}

int upload_padded(exc_index* ix, const void* src, size_t bytes, const void** out)
{
    DevBuf b;
    size_t total = bytes + kPadWords * 4;
    if (int rc = dev_alloc(ix->ctx, b, total)) return rc;
    ix->bufs.push_back(b);
    ix->bytes += b.bytes;
    if (bytes > kStreamChunk) {
        if (int rc = stream_to_device(ix->ctx, b.p, src, bytes)) return rc;
    } else if (bytes) {
        HIP_TRY(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
    }
    HIP_TRY(hipMemset(static_cast<char*>(b.p) + bytes, 0, total - bytes));
    *out = b.p;
    return EXC_OK;
}

uint32_t block_need_words(const uint32_t* blk, uint32_t len)
{
    uint32_t d = blk[0];
    uint32_t lb = desc_lb(d), eb = desc_eb(d), hb = desc_hb(d), en = desc_en(d);
    return 1 + words_for(len, lb) + (en ? words_for(en, eb) + words_for(en, hb) : 0);
}

// The reference's per-bucket HS_SegLRC arrays from the device's compact
// layout (records of non-empty buckets, occupancy tables): every bucket's
// record, an empty one as the host builder writes it (fit {1, 0, 0, 0}, base
// 0, the offset and first element of the next non-empty bucket, or the
// list's block words and length after the last), the tables
// [shift][nb][start_0 .. start_nb] and terms that point at them.
void hs_expand(const std::vector<TermRec>& terms, const std::vector<SegRec>& crec, const std::vector<uint32_t>& occ,
               const std::vector<uint32_t>& words, std::vector<TermRec>& fterms, std::vector<SegRec>& frec,
               std::vector<uint32_t>& table)
{
    fterms = terms;
    frec.clear();
    table.clear();
    for (size_t t = 0; t < terms.size(); ++t) {
        const TermRec& R = terms[t];
        const bool dense = R.aux == kDenseBuckets;
        const uint32_t nb = dense ? R.nseg : occ[R.aux];
        const uint64_t first = frec.size();
        auto present = [&](uint32_t b) { return dense || ((occ[R.aux + 2 + 2 * (b / 32)] >> (b % 32)) & 1u); };
        uint32_t r = 0;
        for (uint32_t b = 0; b < nb; ++b) frec.push_back(present(b) ? crec[R.segnum + r++] : SegRec{});
        uint32_t next_off = 0, next_head = R.n;
        if (R.nseg) {  // after the last block: the list's block words
            const SegRec& q = crec[R.segnum + R.nseg - 1];
            next_off = q.off + block_need_words(words.data() + R.woff + q.off, R.n - q.head);
        }
        for (uint32_t b = nb; b-- > 0;) {
            SegRec& q = frec[first + b];
            if (present(b)) {
                next_off = q.off;
                next_head = q.head;
            } else {
                q = SegRec{next_off, 0u, 1.0f, 0.0f, 0u, 0u, 0, next_head};
            }
        }
        TermRec& F = fterms[t];
        F.segnum = (uint32_t)first;
        F.nseg = nb;
        F.aux = (uint32_t)table.size() + 2;
        table.push_back((uint32_t)R.icpt);
        table.push_back(nb);
        for (uint32_t b = 0; b < nb; ++b) table.push_back(frec[first + b].head);
        table.push_back(R.n);
    }
}

// Words of a NewPFor block of len values whose words are blk[0 .. avail), or 0
// if it is malformed: its Simple16 stream must end inside avail, its
// exception positions must lie inside the block and its descriptor must have
// no bits between en and b.  Reads only words below avail; all is scratch.
uint32_t npfd_block_words(const uint32_t* blk, uint32_t len, uint64_t avail, std::vector<uint32_t>& all)
{
    const uint32_t d = blk[0], b = desc_lb(d), en = desc_en(d);
    if ((d >> 16) & 1023u || en > len) return 0;
    uint64_t w = 1 + (uint64_t)words_for(len, b);
    if (w > avail) return 0;
    if (!en) return (uint32_t)w;
    if (all.size() < 2 * (size_t)en + 28) all.resize(2 * (size_t)en + 28);
    for (uint32_t got = 0; got < 2 * en; ++w) {
        if (w >= avail) return 0;
        const uint32_t c = s16_count(blk[w]);
        s16_unpack(blk[w], all.data() + got, c);
        got += c;
    }
    uint64_t pos = all[0];
    for (uint32_t i = 0; i < en; ++i) {
        if (pos >= len) return 0;
        if (i + 1 < en) pos += (uint64_t)all[i + 1] + 1;
    }
    return (uint32_t)w;
}

}  // namespace

// ---------------------------------------------------------------- context

namespace {
void noalloc_setup(int device);  // exc_intersect_noalloc's scratch address, once per device
}  // namespace

extern "C" int exc_ctx_create(int device, uint64_t budget, exc_ctx** out)
{
    if (!out) return set_error(EXC_ERR_ARG, "exc_ctx_create: null out");
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return set_error(EXC_ERR_ARG, "exc_ctx_create: no device %d", device);
    HIP_TRY(hipSetDevice(device));
    auto* c = new exc_ctx();
    c->device = device;
    c->budget = budget;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->cus = prop.multiProcessorCount;
    if (c->cus <= 0) c->cus = 256;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return set_error(EXC_ERR_DEVICE, "hipStreamCreate failed");
    }
    for (auto& e : c->ev) (void)hipEventCreate(&e);
    for (auto& e : c->pev) (void)hipEventCreateWithFlags(&e, hipEventDisableSystemFence);
    noalloc_setup(device);  // the allocation-free entry point's scratch address on this device
    *out = c;
    return EXC_OK;
}

extern "C" void exc_ctx_destroy(exc_ctx* c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    drop_graph(c);
    for (DevBuf* b : {&c->d_queries, &c->d_qrecs, &c->d_units, &c->d_launch, &c->d_stage, &c->d_ucount, &c->d_uoff,
                      &c->d_bsum, &c->d_result, &c->d_qcum, &c->d_counter})
        dev_free(c, *b);
    for (auto& kv : c->user_bufs) (void)hipFree(kv.first);
    for (auto& kv : c->host_bufs) (void)hipHostFree(kv.first);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : c->pev)
        if (e) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" int exc_device_alloc(exc_ctx* c, uint64_t bytes, void** d_ptr)
{
    if (!c || !d_ptr || !bytes) return set_error(EXC_ERR_ARG, "exc_device_alloc: bad argument");
    HIP_TRY(hipSetDevice(c->device));
    if (c->budget && c->used + bytes > c->budget)
        return set_error(EXC_ERR_BUDGET, "device-memory budget of %llu bytes exceeded", (unsigned long long)c->budget);
    void* p = nullptr;
    HIP_TRY(hipMalloc(&p, bytes));
    c->user_bufs[p] = bytes;
    c->used += bytes;
    *d_ptr = p;
    return EXC_OK;
}

extern "C" int exc_device_free(exc_ctx* c, void* d_ptr)
{
    if (!c) return set_error(EXC_ERR_ARG, "exc_device_free: null ctx");
    auto it = c->user_bufs.find(d_ptr);
    if (it == c->user_bufs.end()) return set_error(EXC_ERR_ARG, "exc_device_free: not a buffer of this context");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    (void)hipFree(d_ptr);
    c->used -= it->second;
    c->user_bufs.erase(it);
    return EXC_OK;
}

extern "C" int exc_host_alloc(exc_ctx* c, uint64_t bytes, void** h_ptr)
{
    if (!c || !h_ptr || !bytes) return set_error(EXC_ERR_ARG, "exc_host_alloc: bad argument");
    HIP_TRY(hipSetDevice(c->device));
    void* p = nullptr;
    HIP_TRY(hipHostMalloc(&p, bytes, hipHostMallocDefault));
    c->host_bufs[p] = bytes;
    *h_ptr = p;
    return EXC_OK;
}

extern "C" int exc_host_free(exc_ctx* c, void* h_ptr)
{
    if (!c || !h_ptr) return set_error(EXC_ERR_ARG, "exc_host_free: bad argument");
    auto it = c->host_bufs.find(h_ptr);
    if (it == c->host_bufs.end()) return set_error(EXC_ERR_ARG, "exc_host_free: not a buffer of this context");
    HIP_TRY(hipHostFree(h_ptr));
    c->host_bufs.erase(it);
    return EXC_OK;
}

extern "C" int exc_memcpy(exc_ctx* c, void* dst, const void* src, uint64_t bytes, int to_device)
{
    if (!c || (bytes && (!dst || !src))) return set_error(EXC_ERR_ARG, "exc_memcpy: bad argument");
    if (!bytes) return EXC_OK;
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(dst, src, bytes, to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost));
    return EXC_OK;
}

extern "C" int exc_ctx_set_option(exc_ctx* c, int key, int64_t v)
{
    if (!c) return set_error(EXC_ERR_ARG, "exc_ctx_set_option: null ctx");
    switch (key) {
    case EXC_OPT_SEARCH:
        if (v < 0 || v > 3) return set_error(EXC_ERR_ARG, "search mode must be 0, 1, 2 or 3");
        c->search = (int)v;
        ++c->opt_gen;
        c->tuned = false;  // an explicit knob replaces the tuned table
        break;
    case EXC_OPT_TPE:
        if (v != 0 && v != 1 && v != 64) return set_error(EXC_ERR_ARG, "tpe must be 0, 1 or 64");
        c->tpe = (int)v;
        ++c->opt_gen;
        c->tuned = false;  // an explicit knob replaces the tuned table
        break;
    case EXC_OPT_PERSISTENT: c->persistent = v != 0; break;
    case EXC_OPT_GRAPH: c->graph = v != 0; break;
    case EXC_OPT_TIMING: c->timing = v != 0; break;
    case EXC_OPT_COOP_MAX:
        if (v < 0 || v > 64) return set_error(EXC_ERR_ARG, "coop_max must be in [0, 64]");
        c->coop_max = (int)v;
        ++c->opt_gen;
        c->tuned = false;  // an explicit knob replaces the tuned table
        break;
    default: return set_error(EXC_ERR_ARG, "unknown option %d", key);
    }
    drop_graph(c);
    return EXC_OK;
}

extern "C" int exc_ctx_last_times(exc_ctx* c, float* probe_ms, float* invoke_ms)
{
    if (!c) return set_error(EXC_ERR_ARG, "null ctx");
    if (probe_ms) *probe_ms = c->probe_ms;
    if (invoke_ms) *invoke_ms = c->invoke_ms;
    return EXC_OK;
}

// ---------------------------------------------------------------- index

namespace {
int fill_pfd_anchors(exc_ctx* c, exc_index* ix);
}

extern "C" int exc_index_upload(exc_ctx* c, const exc_index_host* h, exc_index** out)
{
    if (!c || !h || !out) return set_error(EXC_ERR_ARG, "exc_index_upload: null argument");
    const int codec = h->codec;
    if (codec < EXC_LRC || codec > EXC_RAW_HS) return set_error(EXC_ERR_ARG, "unknown codec %d", codec);
    const bool raw = codec == EXC_RAW || codec == EXC_RAW_HS;
    const bool raw_hs = codec == EXC_RAW_HS;
    const uint32_t seg = raw ? 128u : h->seg_size;  // raw lists: the unit planner's block size
    const bool hs = codec == EXC_HS_SEGLRC;
    if (hs || raw) {
        // the HS kernels walk the shortest list in chunks of 128 elements
    } else if (seg != 128 && seg != 256) {
        return set_error(EXC_ERR_ARG, "segment size %u unsupported (128 or 256)", seg);
    }
    const uint32_t nt = h->nterms;
    if (nt && (!h->ind1 || !h->ind2)) return set_error(EXC_ERR_FORMAT, "index is missing .ind1/.ind2");
    if (nt && !raw && (!h->seg_num || !h->seg_offset))
        return set_error(EXC_ERR_FORMAT, "index is missing .segNum/.segOffset");
    if (!hs && !raw && !h->seg_head) return set_error(EXC_ERR_FORMAT, "index is missing .segHead");
    if (codec == EXC_LRC && !h->lrc_info) return set_error(EXC_ERR_FORMAT, "LRC index is missing .regression_info");
    if ((codec == EXC_SEGLRC || hs) && (!h->seg_info || !h->seg_base))
        return set_error(EXC_ERR_FORMAT, "index is missing .regression_info/.ind2_distance_base");
    if ((hs || raw_hs) && nt && (!h->hash_offset || !h->hash_info))
        return set_error(EXC_ERR_FORMAT, "HS index is missing .HashSeg_offset/.HashSeg_info");

    // ---- validate the layout so that every device read stays inside its array
    std::vector<TermRec> recs(nt);
    std::vector<uint32_t> lwords(nt, 0);
    std::vector<uint32_t> s16_scratch;  // NewPFor exception streams, decoded for the check
    const uint64_t W = h->ind2_words;
    for (uint32_t t = 0; t < nt; ++t) {
        const exc_term_ind1& r = h->ind1[t];
        if (r.m_off % 4) return set_error(EXC_ERR_FORMAT, "term %u: .ind1 offset not word aligned", t);
        const uint64_t w0 = r.m_off / 4;
        const uint64_t w1 = (t + 1 < nt) ? h->ind1[t + 1].m_off / 4 : W;
        if (w0 > w1 || w1 > W) return set_error(EXC_ERR_FORMAT, "term %u: list words outside .ind2", t);
        lwords[t] = (uint32_t)std::min<uint64_t>(w1 - w0, 0xffffffffull);
        const uint32_t n = r.m_urlcount;
        TermRec& R = recs[t];
        std::memset(&R, 0, sizeof R);
        R.woff = w0;
        R.n = n;
        R.segnum = raw ? 0u : h->seg_num[t];
        R.model = kNoModel;
        R.term = t;
        if (raw) {
            // one word per docID, strictly increasing, below the kernels' empty-slot value
            if (w1 - w0 < n) return set_error(EXC_ERR_FORMAT, "term %u: list words outside .ind2", t);
            const uint32_t* L = h->ind2 + w0;
            for (uint32_t i = 0; i < n; ++i)
                if ((i && L[i] <= L[i - 1]) || L[i] == 0xffffffffu)
                    return set_error(EXC_ERR_FORMAT, "term %u: raw list not strictly increasing at %u", t, i);
            R.nseg = (n + seg - 1) / seg;
            R.slope = 1.0f;
            R.range_l = R.range_r = 4.0e9f;
            if (codec == EXC_RAW && h->gu_regression && n > 1 && h->gu_regression[t].fSlope > 0.0f) {
                // the list's regression line and ranges: x = (v - icpt) / slope,
                // every element i with x - nRangeLeft < i < x + nRangeRight
                const exc_gu_regression_info& g = h->gu_regression[t];
                R.slope = g.fSlope;
                std::memcpy(&R.icpt, &g.fIntercept, 4);
                R.range_l = (float)g.nRangeLeft;
                R.range_r = (float)g.nRangeRight;
            }
            if (raw_hs) {
                // [shift][start_0 .. start_nb]: nb is where the starts reach n
                // (the last bucket holds the list's maximum, so no earlier
                // start equals n); starts ascend from 0, and every bucket
                // brackets its docIDs, so a probe that searches bucket
                // min(v >> shift, nb - 1) is exact and reads inside the list
                const uint64_t ho = h->hash_offset[t];
                if (ho + 3 > h->hash_words) return set_error(EXC_ERR_FORMAT, "term %u: hash table outside _HashSeg_info", t);
                const uint32_t* ht = h->hash_info + ho;
                const uint32_t shift = ht[0];
                if (shift > 31 || ht[1] != 0) return set_error(EXC_ERR_FORMAT, "term %u: bad hash table header", t);
                uint64_t nb = 1;
                while (ht[1 + nb] != n) {
                    if (ht[1 + nb] < ht[nb] || ht[1 + nb] > n || ho + 2 + nb >= h->hash_words)
                        return set_error(EXC_ERR_FORMAT, "term %u: bucket starts do not reach the list length", t);
                    ++nb;
                }
                if (nb > 0xffffffffull) return set_error(EXC_ERR_FORMAT, "term %u: too many buckets", t);
                for (uint64_t b = 1; b < nb; ++b) {
                    const uint32_t s = ht[1 + b];
                    if ((s > 0 && (L[s - 1] >> shift) >= b) || (s < n && (L[s] >> shift) < b))
                        return set_error(EXC_ERR_FORMAT, "term %u: bucket %llu does not bracket its docIDs", t,
                                         (unsigned long long)b);
                }
                R.nseg = (uint32_t)nb;
                R.aux = (uint32_t)(ho + 1);
                R.icpt = (int32_t)shift;
            }
            continue;
        }
        uint32_t nseg;
        std::vector<uint32_t> blen;
        if (hs) {
            const uint64_t ho = h->hash_offset[t];
            if (ho + 3 > h->hash_words) return set_error(EXC_ERR_FORMAT, "term %u: hash table outside .HashSeg_info", t);
            const uint32_t* ht = h->hash_info + ho;
            nseg = ht[1];
            if (nseg == 0 || ho + 3 + (uint64_t)nseg > h->hash_words + 1 || ht[0] > 31)
                return set_error(EXC_ERR_FORMAT, "term %u: bad hash table header", t);
            if (ho + 2 + (uint64_t)nseg >= h->hash_words) return set_error(EXC_ERR_FORMAT, "term %u: hash table truncated", t);
            if (ht[2] != 0 || ht[2 + nseg] != n) return set_error(EXC_ERR_FORMAT, "term %u: bucket starts do not span the list", t);
            for (uint32_t b = 0; b < nseg; ++b) {
                if (ht[3 + b] < ht[2 + b]) return set_error(EXC_ERR_FORMAT, "term %u: bucket starts decrease", t);
                blen.push_back(ht[3 + b] - ht[2 + b]);
            }
            R.aux = (uint32_t)(ho + 2);
            R.icpt = (int32_t)ht[0];
        } else {
            nseg = (n + seg - 1) / seg;
            for (uint32_t s = 0; s < nseg; ++s) blen.push_back(std::min(seg, n - s * seg));
        }
        R.nseg = nseg;
        if ((uint64_t)R.segnum + nseg > h->nseg) return set_error(EXC_ERR_FORMAT, "term %u: segments outside .segOffset", t);
        const uint64_t lw = w1 - w0;
        for (uint32_t s = 0; s < nseg; ++s) {
            const uint64_t o = h->seg_offset[R.segnum + s];
            if (blen[s] == 0) { if (o > lw) return set_error(EXC_ERR_FORMAT, "term %u: bad offset of empty bucket", t); continue; }
            if (o >= lw) return set_error(EXC_ERR_FORMAT, "term %u segment %u: offset outside the list", t, s);
            const uint32_t* blk = h->ind2 + w0 + o;
            if (codec == EXC_NEWPFD) {
                if (!npfd_block_words(blk, blen[s], lw - o, s16_scratch))
                    return set_error(EXC_ERR_FORMAT, "term %u segment %u: malformed NewPFor block", t, s);
                continue;
            }
            uint32_t need = block_need_words(blk, blen[s]);
            if (o + need > lw) return set_error(EXC_ERR_FORMAT, "term %u segment %u: block overruns the list", t, s);
            if (codec != EXC_PARAPFD && desc_en(blk[0]))
                return set_error(EXC_ERR_FORMAT,
                                 "term %u segment %u: LRC-family blocks must have no exceptions (FRAC = 0)", t, s);
        }
        if (codec == EXC_LRC) {
            R.slope = h->lrc_info[t].fSlope;
            R.icpt = h->lrc_info[t].iIntercept;
            R.range_l = R.range_r = h->lrc_info[t].fSearchRange;
        } else if (codec == EXC_SEGLRC) {
            exc_seg_info li{1.0f, 0.0f, 0xffffffffu, 0xffffffffu};
            if (h->list_info) li = h->list_info[t];
            R.slope = li.fSlope;
            std::memcpy(&R.icpt, &li.fIntercept, 4);
            R.range_l = (float)li.nRangeLeft;
            R.range_r = (float)li.nRangeRight;
            if (!h->list_info || !(li.fSlope > 0.0f)) { R.slope = 1.0f; R.range_l = R.range_r = 4.0e9f; }
        }
        if (codec == EXC_LRC && !(R.slope > 0.0f)) { R.slope = 1.0f; R.range_l = R.range_r = 4.0e9f; }
    }

    // one 32-byte record per segment / non-empty bucket: offset, descriptor,
    // fit, base, head; HS lists with empty buckets get an occupancy table
    std::vector<SegRec> srec(raw || hs ? 0 : h->nseg);
    std::vector<uint32_t> chead, occ;
    for (uint32_t t = 0; t < nt && hs; ++t) {
        TermRec& R = recs[t];
        const uint32_t* ht = h->hash_info + h->hash_offset[t] + 2;
        const uint32_t nb = R.nseg;
        const uint64_t first = srec.size();
        std::vector<uint32_t> bits((nb + 31) / 32, 0u);
        for (uint32_t b = 0; b < nb; ++b) {
            if (ht[b + 1] == ht[b]) continue;
            bits[b / 32] |= 1u << (b % 32);
            const uint64_t g = (uint64_t)R.segnum + b;
            SegRec q;
            q.off = h->seg_offset[g];
            q.desc = h->ind2[R.woff + q.off];
            q.slope = h->seg_info[g].fSlope;
            q.icpt = h->seg_info[g].fIntercept;
            q.rl = h->seg_info[g].nRangeLeft;
            q.rr = h->seg_info[g].nRangeRight;
            q.base = h->seg_base[g];
            if (!(q.slope > 0.0f)) { q.slope = 1.0f; q.rl = q.rr = 0x7fffffffu; }
            q.head = ht[b];
            srec.push_back(q);
            chead.push_back(ht[b]);
        }
        const uint64_t nr = srec.size() - first;
        if (srec.size() > 0xffffffffull || occ.size() + hs_occ_words(nb) > 0xfffffffeull)
            return set_error(EXC_ERR_CAPACITY, "HS index: more than 2^32 records or table words");
        R.segnum = (uint32_t)first;
        R.nseg = (uint32_t)nr;
        R.aux = kDenseBuckets;
        if (nr == nb) continue;
        R.aux = (uint32_t)occ.size();
        occ.push_back(nb);
        occ.push_back(0);
        uint32_t before = 0;
        for (uint32_t w : bits) {
            occ.push_back(w);
            occ.push_back(before);
            before += (uint32_t)__builtin_popcount(w);
        }
    }
    for (uint32_t t = 0; t < nt && !raw && !hs; ++t) {  // raw lists have no block records
        const TermRec& R = recs[t];
        for (uint32_t s = 0; s < R.nseg; ++s) {
            const uint64_t g = (uint64_t)R.segnum + s;
            SegRec& q = srec[g];
            q.off = h->seg_offset[g];
            q.desc = h->ind2[R.woff + q.off];
            q.slope = 1.0f;
            q.icpt = 0.0f;
            q.rl = q.rr = 0;
            q.base = 0;
            if (codec == EXC_SEGLRC) {
                q.slope = h->seg_info[g].fSlope;
                q.icpt = h->seg_info[g].fIntercept;
                q.rl = h->seg_info[g].nRangeLeft;
                q.rr = h->seg_info[g].nRangeRight;
                q.base = h->seg_base[g];
                if (!(q.slope > 0.0f)) { q.slope = 1.0f; q.rl = q.rr = 0x7fffffffu; }
            }
            q.head = h->seg_head[g];
            // ParaPFD keeps the block's .segMedian row in rl (no fit ranges)
            if (codec == EXC_PARAPFD && h->seg_median) q.rl = h->seg_median[g];
        }
        if (codec == EXC_PARAPFD) {
            float icpt;
            heads_line(h->seg_head + R.segnum, R.nseg, seg, recs[t].slope, icpt, recs[t].range_l, recs[t].range_r);
            std::memcpy(&recs[t].icpt, &icpt, 4);
        }
    }

    auto* ix = new exc_index();
    ix->ctx = c;
    ix->codec = codec;
    ix->seg = hs ? 128 : seg;
    ix->nterms = nt;
    ix->nseg = hs ? srec.size() : h->nseg;
    ix->nwords = W;
    ix->hash_words = hs ? occ.size() : raw_hs ? h->hash_words : 0;
    ix->n.resize(nt);
    for (uint32_t t = 0; t < nt; ++t) ix->n[t] = recs[t].n;
    HIP_TRY(hipSetDevice(c->device));
    int rc = EXC_OK;
    const void* p;
    auto up = [&](const void* src, size_t bytes) -> const void* {
        if (rc) return nullptr;
        rc = upload_padded(ix, src, bytes, &p);
        return rc ? nullptr : p;
    };
    ix->dix.codec = codec;
    ix->dix.seg = ix->seg;
    ix->dix.flags = (codec == EXC_PARAPFD && h->seg_median) ? kIxMedians : 0u;
    ix->dix.words = (const uint32_t*)up(h->ind2, W * 4);
    ix->dix.terms = (const TermRec*)up(recs.data(), recs.size() * sizeof(TermRec));
    ix->d_terms = const_cast<TermRec*>(ix->dix.terms);
    ix->recs = recs;
    ix->lwords = lwords;
    ix->dix.rec = (const SegRec*)up(srec.data(), srec.size() * sizeof(SegRec));
    if (hs) {
        ix->dix.seg_head = (const uint32_t*)up(chead.data(), chead.size() * 4);
        ix->dix.hash = (const uint32_t*)up(occ.data(), occ.size() * 4);
    } else {
        ix->dix.seg_head = (const uint32_t*)up(h->seg_head, h->nseg * 4);
        if (raw_hs) ix->dix.hash = (const uint32_t*)up(h->hash_info, h->hash_words * 4);
    }
    if (rc) {
        exc_index_free(ix);
        return rc;
    }
    // ParaPFD-128: each block's quarter values for the sparse probe (without
    // room in the context's budget the probe keeps its other paths)
    if ((rc = fill_pfd_anchors(c, ix)) && rc != EXC_ERR_BUDGET) {
        exc_index_free(ix);
        return rc;
    }
    *out = ix;
    return EXC_OK;
}

extern "C" void exc_index_free(exc_index* ix)
{
    if (!ix) return;
    if (ix->ctx) {
        (void)hipStreamSynchronize(ix->ctx->stream);
        if (ix->ctx->ix == ix) {
            ix->ctx->ix = nullptr;
            ix->ctx->state = 0;
            drop_graph(ix->ctx);
        }
        for (auto& b : ix->bufs) dev_free(ix->ctx, b);
        dev_free(ix->ctx, ix->d_models);
        dev_free(ix->ctx, ix->d_params);
    }
    delete ix;
}

extern "C" uint64_t exc_index_device_bytes(const exc_index* ix) { return ix ? ix->bytes : 0; }

extern "C" int exc_index_lengths(const exc_index* ix, uint32_t* out, uint32_t cap, uint32_t* nterms)
{
    if (!ix || !nterms) return set_error(EXC_ERR_ARG, "exc_index_lengths: null argument");
    *nterms = ix->nterms;
    if (!out) return EXC_OK;
    if (cap < ix->nterms) return set_error(EXC_ERR_CAPACITY, "exc_index_lengths: %u lists, room for %u", ix->nterms, cap);
    std::copy(ix->n.begin(), ix->n.end(), out);
    return EXC_OK;
}

// A read-only private mapping of one index file: nothing is read until the
// validation or the staged upload touches it, and a 40 GB .ind2 never needs
// 40 GB of heap.
struct MappedFile {
    const char* p = nullptr;
    size_t n = 0;
    MappedFile() = default;
    MappedFile(const MappedFile&) = delete;
    MappedFile& operator=(const MappedFile&) = delete;
    ~MappedFile()
    {
        if (p) munmap(const_cast<char*>(p), n);
    }
    const char* data() const { return p; }
    size_t size() const { return n; }
    bool empty() const { return n == 0; }
};

static int map_file(const std::string& path, MappedFile& out, bool optional)
{
    const int fd = open(path.c_str(), O_RDONLY);
    if (fd < 0) {
        if (optional) return EXC_OK;
        return set_error(EXC_ERR_IO, "cannot open %s", path.c_str());
    }
    struct stat st;
    if (fstat(fd, &st) != 0) {
        close(fd);
        return set_error(EXC_ERR_IO, "cannot stat %s", path.c_str());
    }
    out.n = (size_t)st.st_size;
    if (out.n) {
        void* m = mmap(nullptr, out.n, PROT_READ, MAP_PRIVATE, fd, 0);
        if (m == MAP_FAILED) {
            close(fd);
            out.n = 0;
            return set_error(EXC_ERR_IO, "cannot map %s", path.c_str());
        }
        (void)madvise(m, out.n, MADV_SEQUENTIAL);
        out.p = static_cast<const char*>(m);
    }
    close(fd);
    return EXC_OK;
}

extern "C" int exc_index_load_files(exc_ctx* c, int codec, uint32_t seg, const char* dir, const char* name,
                                    exc_index** out)
{
    if (!c || !dir || !name || !out) return set_error(EXC_ERR_ARG, "exc_index_load_files: null argument");
    std::string p = std::string(dir) + "/" + name;
    MappedFile ind1, ind2, segnum, segoff, head, median, reg, regl, base, hoff, hinfo;
    int rc;
    const bool hs = codec == EXC_HS_SEGLRC;
    if ((rc = map_file(p + ".ind1", ind1, false)) || (rc = map_file(p + ".ind2", ind2, false))) return rc;
    if (ind1.size() % sizeof(exc_term_ind1)) return set_error(EXC_ERR_FORMAT, ".ind1 size is not a multiple of 16");
    if (codec == EXC_RAW || codec == EXC_RAW_HS) {
        // the uncompressed programs' index: .ind1 + the docIDs in .ind2, and for
        // the hash-segment program its tables (cuda_hashSegment.cu:157-186)
        exc_index_host h;
        std::memset(&h, 0, sizeof h);
        h.codec = codec;
        h.seg_size = 128;
        h.nterms = (uint32_t)(ind1.size() / sizeof(exc_term_ind1));
        h.ind1 = (const exc_term_ind1*)ind1.data();
        h.ind2 = (const uint32_t*)ind2.data();
        h.ind2_words = ind2.size() / 4;
        if (codec == EXC_RAW) {  // GU/cuda_runRegressionAlgorithm's at_regression_info, when present
            if ((rc = map_file(p + "_regression_info", reg, true))) return rc;
            if (!reg.empty()) {
                if (reg.size() != (size_t)h.nterms * sizeof(exc_gu_regression_info))
                    return set_error(EXC_ERR_FORMAT, "_regression_info length differs from .ind1");
                h.gu_regression = (const exc_gu_regression_info*)reg.data();
            }
        }
        if (codec == EXC_RAW_HS) {
            if ((rc = map_file(p + "_HashSeg_offset", hoff, false)) || (rc = map_file(p + "_HashSeg_info", hinfo, false)))
                return rc;
            if (hoff.size() / 4 != h.nterms) return set_error(EXC_ERR_FORMAT, "_HashSeg_offset length differs from .ind1");
            h.hash_offset = (const uint32_t*)hoff.data();
            h.hash_info = (const uint32_t*)hinfo.data();
            h.hash_words = hinfo.size() / 4;
        }
        return exc_index_upload(c, &h, out);
    }
    if ((rc = map_file(p + ".segNum", segnum, false)) || (rc = map_file(p + ".segOffset", segoff, false))) return rc;
    if (!hs && (rc = map_file(p + ".segHead", head, false))) return rc;
    if ((rc = map_file(p + ".segMedian", median, true))) return rc;
    const bool gaps = codec == EXC_PARAPFD || codec == EXC_NEWPFD;
    if (!gaps && (rc = map_file(p + ".regression_info", reg, false))) return rc;
    if ((rc = map_file(p + ".regression_info_list", regl, true))) return rc;
    if ((codec == EXC_SEGLRC || hs) && (rc = map_file(p + ".ind2_distance_base", base, false))) return rc;
    if (hs && ((rc = map_file(p + ".HashSeg_offset", hoff, false)) || (rc = map_file(p + ".HashSeg_info", hinfo, false))))
        return rc;
    exc_index_host h;
    std::memset(&h, 0, sizeof h);
    h.codec = codec;
    h.seg_size = seg;
    h.nterms = (uint32_t)(ind1.size() / sizeof(exc_term_ind1));
    h.ind1 = (const exc_term_ind1*)ind1.data();
    h.ind2 = (const uint32_t*)ind2.data();
    h.ind2_words = ind2.size() / 4;
    h.seg_num = (const uint32_t*)segnum.data();
    h.seg_offset = (const uint32_t*)segoff.data();
    h.nseg = segoff.size() / 4;
    if (segnum.size() / 4 != h.nterms) return set_error(EXC_ERR_FORMAT, ".segNum length differs from .ind1");
    if (!hs) {
        if (head.size() != segoff.size()) return set_error(EXC_ERR_FORMAT, ".segHead length differs from .segOffset");
        h.seg_head = (const uint32_t*)head.data();
    }
    if (median.size() == segoff.size() && !median.empty()) h.seg_median = (const uint32_t*)median.data();
    if (codec == EXC_LRC) {
        if (reg.size() != (size_t)h.nterms * sizeof(exc_lrc_info)) return set_error(EXC_ERR_FORMAT, "LRC .regression_info size mismatch");
        h.lrc_info = (const exc_lrc_info*)reg.data();
    } else if (!gaps) {
        if (reg.size() != h.nseg * sizeof(exc_seg_info) || base.size() != h.nseg * 4)
            return set_error(EXC_ERR_FORMAT, ".regression_info / .ind2_distance_base size mismatch");
        h.seg_info = (const exc_seg_info*)reg.data();
        h.seg_base = (const int32_t*)base.data();
        if (regl.size() == (size_t)h.nterms * sizeof(exc_seg_info)) h.list_info = (const exc_seg_info*)regl.data();
    }
    if (hs) {
        if (hoff.size() / 4 != h.nterms) return set_error(EXC_ERR_FORMAT, ".HashSeg_offset length differs from .ind1");
        h.hash_offset = (const uint32_t*)hoff.data();
        h.hash_info = (const uint32_t*)hinfo.data();
        h.hash_words = hinfo.size() / 4;
    }
    return exc_index_upload(c, &h, out);
}

// ---------------------------------------------------------------- planning

namespace {

// Terms ordered by ascending list length with a stable insertion sort
// (LRC_Intersection.cu:157-179), work units = shortest-list segments
// (:182-213; one wave per segment here, one 256-thread block there).
uint32_t ratio_class(const exc_index* ix, const uint32_t* sorted_terms, uint32_t tc)
{
    if (tc < 2) return kRatioClasses - 1;
    const double n0 = std::max<uint32_t>(1, ix->n[sorted_terms[0]]), n1 = ix->n[sorted_terms[1]];
    double r = n1 / n0;
    uint32_t k = 0;
    while (k < 4 && r >= 10.0) { r /= 10.0; ++k; }
    return k;
}

// Knobs of a query of ratio class k: the tuned table, else the context options
// (tpe 1 = one lane per candidate, tpe 64 = always cooperative, tpe 0 =
// cooperative for runs up to coop_max segments; ParaPFD decodes cooperatively).
uint32_t knobs_for(const exc_ctx* c, int codec, uint32_t k)
{
    if (c->tuned) return make_knobs(c->tune.coop_max[k], c->tune.search[k]);
    uint32_t coop = c->tpe == 1 ? 0u : c->tpe == 64 ? 64u : (uint32_t)c->coop_max;
    if (codec == EXC_PARAPFD) coop = 64;
    return make_knobs(coop, (uint32_t)c->search);
}

int plan_queries(exc_ctx* c, exc_index* ix, const uint32_t* qterms, const uint64_t* qoff, uint32_t nq)
{
    if (ix->codec == EXC_NEWPFD)
        return set_error(EXC_ERR_ARG, "NewPFD indexes are for decompression only (as GC/NewPFD in the reference)");
    c->ix = ix;
    c->seg = ix->seg;
    c->queries.resize(nq);
    c->qterms.clear();
    c->qrecs.clear();
    c->units.clear();
    c->short_sum = 0;
    c->nq = nq;
    drop_graph(c);
    uint64_t units = 0;
    std::vector<uint32_t> tmp;
    for (uint32_t q = 0; q < nq; ++q) {
        if (qoff[q + 1] < qoff[q]) return set_error(EXC_ERR_ARG, "query %u: bad offsets", q);
        const uint32_t tc = (uint32_t)(qoff[q + 1] - qoff[q]);
        tmp.assign(qterms + qoff[q], qterms + qoff[q + 1]);
        for (uint32_t t : tmp)
            if (t >= ix->nterms) return set_error(EXC_ERR_ARG, "query %u: term %u out of range", q, t);
        for (uint32_t i = 1; i < tc; ++i) {
            uint32_t x = tmp[i], k = i;
            while (k && ix->n[x] < ix->n[tmp[k - 1]]) { tmp[k] = tmp[k - 1]; --k; }
            tmp[k] = x;
        }
        QueryRec& Q = c->queries[q];
        std::memset(&Q, 0, sizeof Q);
        Q.toff = (uint32_t)c->qterms.size();
        Q.tc = tc;
        Q.first_unit = (uint32_t)units;
        const uint32_t n0 = tc ? ix->n[tmp[0]] : 0;
        Q.nunits = (n0 + kChunk - 1) / kChunk;
        Q.short_off = c->short_sum;
        Q.klass = ratio_class(ix, tmp.data(), tc);
        Q.knobs = knobs_for(c, ix->codec, Q.klass);
        c->short_sum += n0;
        c->qterms.insert(c->qterms.end(), tmp.begin(), tmp.end());
        for (uint32_t t : tmp) c->qrecs.push_back(ix->recs[t]);
        if (tc > 0xffff) return set_error(EXC_ERR_ARG, "query %u: more than 65535 terms", q);
        for (uint32_t k = 0; k < Q.nunits; ++k)
            c->units.push_back(UnitRec{q, k, Q.toff, (uint16_t)tc, (uint16_t)Q.knobs, 0, kUnset,
                                       (uint32_t)c->units.size(), (uint32_t)(Q.short_off + (uint64_t)k * kChunk)});
        units += Q.nunits;
        if (units > 0xffffffffull) return set_error(EXC_ERR_CAPACITY, "more than 2^32 work units in one batch");
    }
    c->nunits = (uint32_t)units;
    c->plan_gen = c->opt_gen;
    c->plan_model_gen = ix->model_gen;
    if (c->short_sum > 0xffffffffull)
        return set_error(EXC_ERR_CAPACITY, "summed shortest-list length %llu exceeds 2^32",
                         (unsigned long long)c->short_sum);
    c->state = 1;
    return EXC_OK;
}

// Launch order of the units: queries by descending |second-shortest| /
// |shortest| (single-term queries last), each query's units together.  The
// high-ratio units are the slowest ones (their probes search, the others
// resolve whole runs through the bitmap), so running them first shortens the
// launch's tail: C2 batch probe 0.653-0.661 ms whatever the callers' query
// order, against 0.675 ms when units ran in the order given.
void build_launch(exc_ctx* c)
{
    const exc_index* ix = c->ix;
    std::vector<uint32_t> order(c->nq);
    std::vector<double> key(c->nq, 0.0);
    for (uint32_t q = 0; q < c->nq; ++q) {
        order[q] = q;
        const QueryRec& Q = c->queries[q];
        if (Q.tc > 1) {
            const double n0 = ix->n[c->qterms[Q.toff]], n1 = ix->n[c->qterms[Q.toff + 1]];
            key[q] = n1 / std::max(1.0, n0);
        }
    }
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return key[a] > key[b]; });
    c->launch.clear();
    c->launch.reserve(c->units.size());
    for (uint32_t q : order) {
        const QueryRec& Q = c->queries[q];
        c->launch.insert(c->launch.end(), c->units.begin() + Q.first_unit, c->units.begin() + Q.first_unit + Q.nunits);
    }
}

int upload_plan(exc_ctx* c, bool sync = true)
{
    if (c->state < 1) return set_error(EXC_ERR_STATE, "htodTransfer before CPUPreprocess");
    HIP_TRY(hipSetDevice(c->device));
    const uint32_t nblk = (c->nunits + kScanBlockItems - 1) / kScanBlockItems;
    int rc;
    if ((rc = dev_alloc(c, c->d_queries, std::max<size_t>(1, c->nq) * sizeof(QueryRec))) ||
        (rc = dev_alloc(c, c->d_qrecs, std::max<size_t>(1, c->qrecs.size()) * sizeof(TermRec))) ||
        (rc = dev_alloc(c, c->d_units, std::max<size_t>(1, c->nunits) * sizeof(UnitRec))) ||
        (rc = dev_alloc(c, c->d_launch, std::max<size_t>(1, c->nunits) * sizeof(UnitRec))) ||
        (rc = dev_alloc(c, c->d_stage, std::max<uint64_t>(1, c->short_sum) * 4)) ||
        (rc = dev_alloc(c, c->d_ucount, std::max<size_t>(1, c->nunits) * 4)) ||
        (rc = dev_alloc(c, c->d_uoff, std::max<size_t>(1, c->nunits) * 4)) ||
        (rc = dev_alloc(c, c->d_bsum, (size_t)(nblk + 1) * 4)) ||
        (rc = dev_alloc(c, c->d_result, std::max<uint64_t>(1, c->short_sum) * 4)) ||
        (rc = dev_alloc(c, c->d_qcum, std::max<size_t>(1, c->nq) * 4)) ||
        (rc = dev_alloc(c, c->d_counter, 256)))
        return rc;
    drop_graph(c);
    // the scan's ticket starts at zero when its buffer is new; every launch
    // leaves it at zero
    if (c->d_counter.p != c->ticket_zeroed) {
        HIP_TRY(hipMemsetAsync(dptr<uint32_t>(c->d_counter) + kScanTicketWord, 0, 4, c->stream));
        c->ticket_zeroed = c->d_counter.p;
    }
    if (c->nq) HIP_TRY(hipMemcpyAsync(c->d_queries.p, c->queries.data(), c->nq * sizeof(QueryRec), hipMemcpyHostToDevice, c->stream));
    if (!c->qrecs.empty())
        HIP_TRY(hipMemcpyAsync(c->d_qrecs.p, c->qrecs.data(), c->qrecs.size() * sizeof(TermRec), hipMemcpyHostToDevice,
                               c->stream));
    build_launch(c);
    if (c->nunits) {
        HIP_TRY(hipMemcpyAsync(c->d_units.p, c->units.data(), (size_t)c->nunits * sizeof(UnitRec),
                               hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(c->d_launch.p, c->launch.data(), (size_t)c->nunits * sizeof(UnitRec),
                               hipMemcpyHostToDevice, c->stream));
    }
    if (sync) HIP_TRY(hipStreamSynchronize(c->stream));
    c->state = 2;
    return EXC_OK;
}

// Options changed after planning: recompute every query's knobs and send the
// query records again (the records live in device memory, so a captured graph
// stays valid).
int refresh_knobs(exc_ctx* c)
{
    if (c->plan_model_gen != c->ix->model_gen) {
        // the index's models changed after planning: fresh record copies, and
        // no graph that captured the old model arrays
        for (size_t i = 0; i < c->qterms.size(); ++i) c->qrecs[i] = c->ix->recs[c->qterms[i]];
        if (!c->qrecs.empty())
            HIP_TRY(hipMemcpyAsync(c->d_qrecs.p, c->qrecs.data(), c->qrecs.size() * sizeof(TermRec),
                                   hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        drop_graph(c);
        c->plan_model_gen = c->ix->model_gen;
    }
    if (c->plan_gen == c->opt_gen) return EXC_OK;
    for (auto& Q : c->queries) Q.knobs = knobs_for(c, c->ix->codec, Q.klass);
    for (auto& U : c->units) U.knobs = (uint16_t)c->queries[U.q].knobs;
    for (auto& U : c->launch) U.knobs = (uint16_t)c->queries[U.q].knobs;
    if (c->nq)
        HIP_TRY(hipMemcpyAsync(c->d_queries.p, c->queries.data(), c->nq * sizeof(QueryRec), hipMemcpyHostToDevice,
                               c->stream));
    if (c->nunits) {
        HIP_TRY(hipMemcpyAsync(c->d_units.p, c->units.data(), (size_t)c->nunits * sizeof(UnitRec),
                               hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(c->d_launch.p, c->launch.data(), (size_t)c->nunits * sizeof(UnitRec),
                               hipMemcpyHostToDevice, c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->plan_gen = c->opt_gen;
    return EXC_OK;
}

int enqueue_invoke(exc_ctx* c, bool timed)
{
    exc_index* ix = c->ix;
    ProbeArgs a;
    std::memset(&a, 0, sizeof a);
    a.ix = ix->dix;
    a.queries = dptr<QueryRec>(c->d_queries);
    a.qrecs = dptr<TermRec>(c->d_qrecs);
    a.units = dptr<UnitRec>(c->d_launch);
    a.nunits = c->nunits;
    a.persistent = c->persistent;
    a.stage = dptr<uint32_t>(c->d_stage);
    a.ucount = dptr<uint32_t>(c->d_ucount);
    a.counter = dptr<uint32_t>(c->d_counter);
    const uint32_t wpb = probe_waves_per_block();
    uint32_t grid;
    if (c->persistent) {
        uint32_t per = (uint32_t)probe_blocks_per_cu(ix->codec, ix->seg);
        grid = std::max<uint32_t>(1, std::min<uint32_t>((uint32_t)c->cus * per, (c->nunits + wpb - 1) / wpb));
        HIP_TRY(launch_reset(a.counter, c->stream));
    } else {
        grid = std::max<uint32_t>(1, (c->nunits + wpb - 1) / wpb);
    }
    if (c->nunits) HIP_TRY(launch_unit_prologue(a, c->stream));
    if (timed) HIP_TRY(hipEventRecord(c->pev[0], c->stream));
    if (c->nunits) HIP_TRY(launch_probe(a, grid, c->stream));
    if (timed) HIP_TRY(hipEventRecord(c->pev[1], c->stream));
    ScanArgs s;
    std::memset(&s, 0, sizeof s);
    s.queries = a.queries;
    s.units = dptr<UnitRec>(c->d_units);  // query order
    s.ucount = a.ucount;
    s.stage = a.stage;
    s.nunits = c->nunits;
    s.nq = c->nq;
    s.uoff = dptr<uint32_t>(c->d_uoff);
    s.bsum = dptr<uint32_t>(c->d_bsum);
    s.result = dptr<uint32_t>(c->d_result);
    s.qcum = dptr<uint32_t>(c->d_qcum);
    s.ticket = dptr<uint32_t>(c->d_counter) + kScanTicketWord;
    HIP_TRY(launch_scan(s, c->stream));
    return EXC_OK;
}

int invoke(exc_ctx* c)
{
    if (c->state < 2 || !c->ix) return set_error(EXC_ERR_STATE, "kernelInvoke before htodTransfer");
    HIP_TRY(hipSetDevice(c->device));
    if (int rc = refresh_knobs(c)) return rc;
    const bool timed = c->timing && !c->graph;
    if (c->graph) {
        if (!c->gexec) {
            drop_graph(c);
            hipGraph_t g;
            HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            int rc = enqueue_invoke(c, false);
            hipError_t e = hipStreamEndCapture(c->stream, &g);
            if (rc) return rc;
            if (e != hipSuccess) return set_error(EXC_ERR_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
            e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) return set_error(EXC_ERR_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
        }
        if (c->timing) HIP_TRY(hipEventRecord(c->ev[0], c->stream));
        HIP_TRY(hipGraphLaunch(c->gexec, c->stream));
        if (c->timing) HIP_TRY(hipEventRecord(c->ev[3], c->stream));
    } else {
        if (timed) HIP_TRY(hipEventRecord(c->ev[0], c->stream));
        if (int rc = enqueue_invoke(c, timed)) return rc;
        if (timed) HIP_TRY(hipEventRecord(c->ev[3], c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->probe_ms = c->invoke_ms = 0;
    if (timed) {
        (void)hipEventElapsedTime(&c->probe_ms, c->pev[0], c->pev[1]);
        (void)hipEventElapsedTime(&c->invoke_ms, c->ev[0], c->ev[3]);
    } else if (c->timing) {
        (void)hipEventElapsedTime(&c->invoke_ms, c->ev[0], c->ev[3]);
    }
    c->state = 3;
    return EXC_OK;
}

int fetch(exc_ctx* c, uint32_t* cum, uint32_t* res, uint64_t cap, uint64_t* nres)
{
    if (c->state < 3) return set_error(EXC_ERR_STATE, "dtohTransfer before kernelInvoke");
    HIP_TRY(hipSetDevice(c->device));
    uint64_t total = 0;
    if (c->nq) {
        HIP_TRY(hipMemcpy(cum, c->d_qcum.p, c->nq * 4, hipMemcpyDeviceToHost));
        total = cum[c->nq - 1];
    }
    if (nres) *nres = total;
    if (total > cap) return set_error(EXC_ERR_CAPACITY, "result buffer holds %llu, need %llu",
                                      (unsigned long long)cap, (unsigned long long)total);
    if (total) HIP_TRY(hipMemcpy(res, c->d_result.p, total * 4, hipMemcpyDeviceToHost));
    return EXC_OK;
}

}  // namespace

// ---------------------------------------------------------------- stages

extern "C" int exc_CPUPreprocess(exc_ctx* c, exc_index* ix, const uint8_t** ptr_in, const uint8_t* ptr_end,
                                 uint64_t threshold, exc_batchInfo* bi)
{
    if (!c || !ix || !ptr_in || !*ptr_in || !ptr_end || !bi) return set_error(EXC_ERR_ARG, "exc_CPUPreprocess: null argument");
    const uint8_t* p = *ptr_in;
    std::vector<uint32_t> terms;
    std::vector<uint64_t> off{0};
    uint64_t sum = 0;
    // LRC_Intersection.cu:140: keep adding queries while the running shortest
    // sum is <= threshold (the query that crosses it joins the batch).
    while (sum <= threshold && p < ptr_end) {
        if (ptr_end - p < 4) return set_error(EXC_ERR_ARG, "truncated query header");
        uint32_t tn;
        std::memcpy(&tn, p, 4);
        p += 4;
        if ((uint64_t)(ptr_end - p) < (uint64_t)tn * 4) return set_error(EXC_ERR_ARG, "truncated query terms");
        uint32_t mn = 0xffffffffu;
        for (uint32_t i = 0; i < tn; ++i) {
            uint32_t t;
            std::memcpy(&t, p + 4 * i, 4);
            if (t >= ix->nterms) return set_error(EXC_ERR_ARG, "query term %u out of range", t);
            terms.push_back(t);
            mn = std::min(mn, ix->n[t]);
        }
        p += 4ull * tn;
        off.push_back(terms.size());
        if (tn) sum += mn;
    }
    const uint32_t nq = (uint32_t)(off.size() - 1);
    if (int rc = plan_queries(c, ix, terms.data(), off.data(), nq)) return rc;
    *ptr_in = p;
    bi->blockNum = c->nunits;
    bi->constantUsedInByte = (uint32_t)std::min<uint64_t>(0xffffffffu, nq * sizeof(QueryRec) + c->qrecs.size() * sizeof(TermRec) + (uint64_t)c->nunits * sizeof(UnitRec));
    bi->ucntShortest_sum = (uint32_t)c->short_sum;
    bi->nTotalQueryNum = nq;
    bi->nLinearInfoInByte = (uint32_t)std::min<uint64_t>(0xffffffffu, c->qterms.size() * sizeof(TermRec));
    return EXC_OK;
}

extern "C" int exc_htodTransfer(exc_ctx* c, const exc_batchInfo* bi)
{
    if (!c || !bi) return set_error(EXC_ERR_ARG, "exc_htodTransfer: null argument");
    if (bi->nTotalQueryNum != c->nq) return set_error(EXC_ERR_STATE, "batchInfo does not match the planned batch");
    return upload_plan(c);
}

extern "C" int exc_kernelInvoke(exc_ctx* c, const exc_batchInfo* bi)
{
    if (!c || !bi) return set_error(EXC_ERR_ARG, "exc_kernelInvoke: null argument");
    if (bi->nTotalQueryNum != c->nq) return set_error(EXC_ERR_STATE, "batchInfo does not match the planned batch");
    return invoke(c);
}

extern "C" int exc_dtohTransfer(exc_ctx* c, const exc_batchInfo* bi, uint32_t* cum, uint32_t* res, uint64_t cap)
{
    if (!c || !bi || (!cum && bi->nTotalQueryNum)) return set_error(EXC_ERR_ARG, "exc_dtohTransfer: null argument");
    if (bi->nTotalQueryNum != c->nq) return set_error(EXC_ERR_STATE, "batchInfo does not match the planned batch");
    return fetch(c, cum, res, cap, nullptr);
}

// LRC_Intersection.cu:354-379
extern "C" int exc_verify(const exc_batchInfo* bi, const uint32_t* cum, const uint32_t* res, exc_checkSum* cs)
{
    if (!bi || !cs || (bi->nTotalQueryNum && !cum)) return set_error(EXC_ERR_ARG, "exc_verify: null argument");
    if (bi->nTotalQueryNum && cum[bi->nTotalQueryNum - 1] && !res)
        return set_error(EXC_ERR_ARG, "exc_verify: null result buffer for %u results", cum[bi->nTotalQueryNum - 1]);
    for (uint32_t i = 0; i < bi->nTotalQueryNum; ++i) {
        const uint32_t b = i ? cum[i - 1] : 0, n = cum[i] - b;
        if (n) {
            cs->checkSum2 += n;
            cs->checkSum3 ^= res[b];
            cs->checkSum3 ^= res[b + n - 1];
            cs->checkSum3 ^= res[b + n / 2];
        } else {
            ++cs->checkSum1;
        }
    }
    return EXC_OK;
}

extern "C" int exc_plan(exc_ctx* c, exc_index* ix, const uint32_t* qterms, const uint64_t* qoff, uint32_t nq,
                        exc_batchInfo* bi)
{
    if (!c || !ix || !qoff || (nq && !qterms && qoff[nq]) || !bi) return set_error(EXC_ERR_ARG, "exc_plan: null argument");
    if (int rc = plan_queries(c, ix, qterms, qoff, nq)) return rc;
    bi->blockNum = c->nunits;
    bi->constantUsedInByte = (uint32_t)std::min<uint64_t>(0xffffffffu, nq * sizeof(QueryRec) + c->qrecs.size() * sizeof(TermRec) + (uint64_t)c->nunits * sizeof(UnitRec));
    bi->ucntShortest_sum = (uint32_t)c->short_sum;
    bi->nTotalQueryNum = nq;
    bi->nLinearInfoInByte = (uint32_t)std::min<uint64_t>(0xffffffffu, c->qterms.size() * sizeof(TermRec));
    return upload_plan(c);
}

extern "C" int exc_intersect(exc_ctx* c, exc_index* ix, const uint32_t* qterms, const uint64_t* qoff, uint32_t nq,
                             uint32_t* cum, uint32_t* res, uint64_t cap, uint64_t* nres)
{
    exc_batchInfo bi;
    if (int rc = exc_plan(c, ix, qterms, qoff, nq, &bi)) return rc;
    if (int rc = invoke(c)) return rc;
    return fetch(c, cum, res, cap, nres);
}

// ---------------------------------------------------------------- allocation-free intersect

// Device scratch of exc_intersect_noalloc: a fixed array in the library's code
// object, one copy per GPU, held until the process ends.  Host code never
// names it in a kernel argument or a copy (the name is a host-side stand-in);
// it uses the address hipGetSymbolAddress returns for the current device.
constexpr size_t kNaScratchBytes = size_t(64) << 20;
static __device__ __align__(256) uint8_t scratch[kNaScratchBytes];

namespace {

constexpr int kNaMaxDevices = 64;
std::mutex g_na_mu;
void* g_na_ptr[kNaMaxDevices] = {};        // scratch's device address on each GPU
hipError_t g_na_err[kNaMaxDevices] = {};   // the lookup's result, once made

// Once per device, with that device current (exc_ctx_create's set-up).
void noalloc_setup(int device)
{
    if (device < 0 || device >= kNaMaxDevices) return;
    std::lock_guard<std::mutex> g(g_na_mu);
    if (g_na_ptr[device] || g_na_err[device] != hipSuccess) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess && p) g_na_ptr[device] = p;
    else g_na_err[device] = e != hipSuccess ? e : hipErrorInvalidSymbol;
}

size_t na_align(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace

// exc_intersect without any allocation: the plan is built on the device from
// the query list copied into the scratch array, and the probe, scan and
// scatter use the scratch for their records, stage and results.  Batches whose
// records would not fit are refused with EXC_ERR_CAPACITY.
extern "C" int exc_intersect_noalloc(exc_ctx* c, exc_index* ix, const uint32_t* qterms, const uint64_t* qoff,
                                     uint32_t nq, uint32_t* cum, uint32_t* res, uint64_t cap, uint64_t* nres)
{
    if (!c || !ix || !qoff || (nq && !cum)) return set_error(EXC_ERR_ARG, "exc_intersect_noalloc: bad argument");
    if (ix->codec == EXC_NEWPFD)
        return set_error(EXC_ERR_ARG, "NewPFD indexes are for decompression only (as GC/NewPFD in the reference)");
    if (c->device < 0 || c->device >= kNaMaxDevices)
        return set_error(EXC_ERR_ARG, "exc_intersect_noalloc: device %d beyond %d", c->device, kNaMaxDevices);
    HIP_TRY(hipSetDevice(c->device));
    noalloc_setup(c->device);  // made by exc_ctx_create already; a no-op then
    uint8_t* base = static_cast<uint8_t*>(g_na_ptr[c->device]);
    if (!base)
        return set_error(EXC_ERR_DEVICE, "scratch address of device %d: %s", c->device,
                         hipGetErrorString(g_na_err[c->device]));
    // sizes, checked against the scratch before anything is sent (no storage:
    // the shortest lengths are summed here and recomputed on the device)
    const uint64_t nterm = nq ? qoff[nq] - qoff[0] : 0;
    if (nterm && !qterms) return set_error(EXC_ERR_ARG, "exc_intersect_noalloc: null terms");
    uint64_t units = 0, short_sum = 0;
    for (uint32_t q = 0; q < nq; ++q) {
        if (qoff[q + 1] < qoff[q]) return set_error(EXC_ERR_ARG, "query %u: bad offsets", q);
        const uint64_t tc = qoff[q + 1] - qoff[q];
        if (tc > kNaMaxTerms)
            return set_error(EXC_ERR_CAPACITY, "query %u: %llu terms, the allocation-free path plans at most %u", q,
                             (unsigned long long)tc, kNaMaxTerms);
        uint32_t n0 = 0;
        for (uint64_t i = 0; i < tc; ++i) {
            const uint32_t t = qterms[qoff[q] + i];
            if (t >= ix->nterms) return set_error(EXC_ERR_ARG, "query %u: term %u out of range", q, t);
            n0 = (i == 0 || ix->n[t] < n0) ? ix->n[t] : n0;
        }
        units += (n0 + kChunk - 1) / kChunk;
        short_sum += n0;
    }
    const uint32_t nblk = (uint32_t)((units + kScanBlockItems - 1) / kScanBlockItems);
    // layout: counters | terms | offsets | queries | term records | units |
    // unit counts | unit offsets | block sums | stage | results | cumulative counts
    size_t o = 0;
    const size_t o_cnt = o;
    o += na_align(256);
    const size_t o_qt = o;
    o += na_align(nterm * 4);
    const size_t o_qo = o;
    o += na_align((nq + 1ull) * 8);
    const size_t o_q = o;
    o += na_align((uint64_t)nq * sizeof(QueryRec));
    const size_t o_tr = o;
    o += na_align(nterm * sizeof(TermRec));
    const size_t o_u = o;
    o += na_align(units * sizeof(UnitRec));
    const size_t o_uc = o;
    o += na_align(units * 4);
    const size_t o_uo = o;
    o += na_align(units * 4);
    const size_t o_bs = o;
    o += na_align((nblk + 1ull) * 4);
    const size_t o_st = o;
    o += na_align(short_sum * 4);
    const size_t o_res = o;
    o += na_align(short_sum * 4);
    const size_t o_cum = o;
    o += na_align((uint64_t)nq * 4);
    if (o > kNaScratchBytes || units > 0xffffffffull)
        return set_error(EXC_ERR_CAPACITY,
                         "batch needs %llu bytes of scratch (%llu shortest-list docIDs, %llu units), the "
                         "allocation-free path has %llu",
                         (unsigned long long)o, (unsigned long long)short_sum, (unsigned long long)units,
                         (unsigned long long)kNaScratchBytes);
    if (nres) *nres = 0;
    if (!nq) return EXC_OK;
    hipStream_t s = c->stream;
    // counters (the persistent queue's claim word and the scan's ticket)
    // start at zero: a fill kernel, not a memset node
    HIP_TRY(launch_zero_words(reinterpret_cast<uint32_t*>(base + o_cnt), 64, s));
    if (nterm) HIP_TRY(hipMemcpyAsync(base + o_qt, qterms + qoff[0], nterm * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(base + o_qo, qoff, (nq + 1ull) * 8, hipMemcpyHostToDevice, s));
    NaPlanArgs pa;
    std::memset(&pa, 0, sizeof pa);
    pa.terms = ix->dix.terms;
    pa.qterms = reinterpret_cast<const uint32_t*>(base + o_qt);
    pa.qoff = reinterpret_cast<const uint64_t*>(base + o_qo);
    pa.nq = nq;
    for (int k = 0; k < kRatioClasses; ++k) pa.knobs[k] = knobs_for(c, ix->codec, (uint32_t)k);
    pa.queries = reinterpret_cast<QueryRec*>(base + o_q);
    pa.qrecs = reinterpret_cast<TermRec*>(base + o_tr);
    pa.units = reinterpret_cast<UnitRec*>(base + o_u);
    HIP_TRY(launch_noalloc_plan(pa, s));
    ProbeArgs a;
    std::memset(&a, 0, sizeof a);
    a.ix = ix->dix;
    a.queries = pa.queries;
    a.qrecs = pa.qrecs;
    a.units = pa.units;  // query order is the launch order here
    a.nunits = (uint32_t)units;
    a.persistent = c->persistent;
    a.stage = reinterpret_cast<uint32_t*>(base + o_st);
    a.ucount = reinterpret_cast<uint32_t*>(base + o_uc);
    a.counter = reinterpret_cast<uint32_t*>(base + o_cnt);
    const uint32_t wpb = probe_waves_per_block();
    uint32_t grid = std::max<uint32_t>(1, (a.nunits + wpb - 1) / wpb);
    if (c->persistent)
        grid = std::max<uint32_t>(
            1, std::min<uint32_t>((uint32_t)c->cus * (uint32_t)probe_blocks_per_cu(ix->codec, ix->seg), grid));
    if (a.nunits) {
        HIP_TRY(launch_unit_prologue(a, s));
        HIP_TRY(launch_probe(a, grid, s));
    }
    ScanArgs sa;
    std::memset(&sa, 0, sizeof sa);
    sa.queries = a.queries;
    sa.units = a.units;
    sa.ucount = a.ucount;
    sa.stage = a.stage;
    sa.nunits = a.nunits;
    sa.nq = nq;
    sa.uoff = reinterpret_cast<uint32_t*>(base + o_uo);
    sa.bsum = reinterpret_cast<uint32_t*>(base + o_bs);
    sa.result = reinterpret_cast<uint32_t*>(base + o_res);
    sa.qcum = reinterpret_cast<uint32_t*>(base + o_cum);
    sa.ticket = a.counter + kScanTicketWord;
    HIP_TRY(launch_scan(sa, s));
    HIP_TRY(hipMemcpyAsync(cum, sa.qcum, (size_t)nq * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const uint64_t total = cum[nq - 1];
    if (nres) *nres = total;
    if (total > cap)
        return set_error(EXC_ERR_CAPACITY, "result buffer holds %llu, need %llu", (unsigned long long)cap,
                         (unsigned long long)total);
    if (total) {
        if (!res) return set_error(EXC_ERR_ARG, "exc_intersect_noalloc: null result buffer");
        HIP_TRY(hipMemcpyAsync(res, sa.result, total * 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    return EXC_OK;
}

// ---------------------------------------------------------------- decompress

namespace {

// Decode whole lists of `terms` into dec (device), concatenated; off[i] =
// offset of terms[i].  Temporary buffers are freed before returning.
// Decode whole lists into device memory: into `dst` when given (capacity
// checked by the caller), else into `dec`, allocated here.  With timing on,
// the kernel's duration lands in probe_ms.
int decode_to_device(exc_ctx* c, exc_index* ix, const std::vector<uint32_t>& terms, DevBuf& dec,
                     std::vector<uint64_t>& off, uint32_t* dst = nullptr)
{
    const uint32_t nt = (uint32_t)terms.size();
    off.assign(nt + 1, 0);
    std::vector<DecUnit> units;
    // units: runs of kDecodeUnit/seg blocks (HS: chunks of seg elements), k = first block
    const uint32_t step = ix->codec == EXC_HS_SEGLRC ? 1u : kDecodeUnit / ix->seg;
    for (uint32_t i = 0; i < nt; ++i) {
        const TermRec& R = ix->recs[terms[i]];
        const uint32_t n = ix->n[terms[i]];
        off[i + 1] = off[i] + n;
        for (uint32_t k = 0; k < (n + ix->seg - 1) / ix->seg; k += step) {
            DecUnit d;
            std::memset(&d, 0, sizeof d);
            d.out = off[i] + (uint64_t)k * ix->seg;
            d.woff = R.woff;
            d.segnum = R.segnum;
            d.n = n;
            d.k = k;
            d.term = terms[i];
            d.slope = R.slope;
            d.icpt = R.icpt;
            d.wend = terms[i] < ix->lwords.size() ? ix->lwords[terms[i]] : kUnset;
            units.push_back(d);
        }
    }
    c->probe_ms = 0;
    if (!off[nt]) return EXC_OK;
    if (units.size() > 0xffffffffull) return set_error(EXC_ERR_CAPACITY, "more than 2^32 decode units");
    DevBuf du;
    int rc;
    if ((!dst && (rc = dev_alloc(c, dec, off[nt] * 4))) || (rc = dev_alloc(c, du, units.size() * sizeof(DecUnit))))
        return rc;
    hipError_t e = hipMemcpy(du.p, units.data(), units.size() * sizeof(DecUnit), hipMemcpyHostToDevice);
    if (e == hipSuccess && c->timing) e = hipEventRecord(c->ev[1], c->stream);
    if (e == hipSuccess)
        e = launch_decompress(ix->dix, dptr<DecUnit>(du), (uint32_t)units.size(), dst ? dst : dptr<uint32_t>(dec),
                              c->stream);
    if (e == hipSuccess && c->timing) e = hipEventRecord(c->ev[2], c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess && c->timing) (void)hipEventElapsedTime(&c->probe_ms, c->ev[1], c->ev[2]);
    dev_free(c, du);
    if (e != hipSuccess) return set_error(EXC_ERR_DEVICE, "decode: %s", hipGetErrorString(e));
    return EXC_OK;
}

// ParaPFD-128 indexes uploaded from host arrays or files: the values at each
// block's quarter positions (kIxQuarters), taken from the lists decoded on the
// device, at most 2^28 docIDs at a time.
int fill_pfd_anchors(exc_ctx* c, exc_index* ix)
{
    if (ix->codec != EXC_PARAPFD || ix->seg != 128 || !(ix->dix.flags & kIxMedians)) return EXC_OK;
    SegRec* rec = const_cast<SegRec*>(ix->dix.rec);
    DevBuf ob;  // the quarters' midpoint values, 4 a block, kept with the index
    if (int rc = dev_alloc(c, ob, std::max<uint64_t>(1, ix->nseg) * 16)) return rc;
    ix->bufs.push_back(ob);
    ix->bytes += ob.bytes;
    uint32_t* oct = static_cast<uint32_t*>(ob.p);
    DevBuf dec, dj;
    std::vector<uint64_t> off;
    std::vector<uint32_t> terms;
    std::vector<AnchorJob> jobs;
    int rc = EXC_OK;
    for (uint32_t t0 = 0; t0 < ix->nterms && !rc;) {
        terms.clear();
        uint64_t tot = 0;
        uint32_t t = t0;
        for (; t < ix->nterms && (terms.empty() || tot + ix->n[t] <= (1ull << 28)); ++t) {
            terms.push_back(t);
            tot += ix->n[t];
        }
        t0 = t;
        if (!tot) continue;
        if ((rc = decode_to_device(c, ix, terms, dec, off))) break;
        jobs.clear();
        for (size_t i = 0; i < terms.size(); ++i) {
            const TermRec& R = ix->recs[terms[i]];
            const uint32_t n = ix->n[terms[i]];
            for (uint32_t s = 0; s < R.nseg; ++s)
                jobs.push_back(AnchorJob{(uint64_t)R.segnum + s, off[i] + (uint64_t)s * 128u,
                                         std::min<uint32_t>(128u, n - s * 128u), 0u});
        }
        if (jobs.size() > 0xffffffffull) {
            rc = set_error(EXC_ERR_CAPACITY, "more than 2^32 blocks in one decoded run");
            break;
        }
        if ((rc = dev_alloc(c, dj, jobs.size() * sizeof(AnchorJob)))) break;
        hipError_t e = hipMemcpy(dj.p, jobs.data(), jobs.size() * sizeof(AnchorJob), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = launch_pfd_anchors(dptr<AnchorJob>(dj), (uint32_t)jobs.size(), dptr<uint32_t>(dec), rec, oct, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = set_error(EXC_ERR_DEVICE, "ParaPFD quarter values: %s", hipGetErrorString(e));
    }
    dev_free(c, dec);
    dev_free(c, dj);
    if (!rc) {
        ix->dix.pfd_oct = oct;
        ix->dix.flags |= kIxQuarters | kIxOctants;
    }
    return rc;
}

}  // namespace

static int check_terms(exc_index* ix, const uint32_t* terms, uint32_t nt, uint64_t* total)
{
    uint64_t n = 0;
    for (uint32_t i = 0; i < nt; ++i) {
        if (terms[i] >= ix->nterms) return set_error(EXC_ERR_ARG, "term %u out of range", terms[i]);
        n += ix->n[terms[i]];
    }
    *total = n;
    return EXC_OK;
}

extern "C" int exc_decompress(exc_ctx* c, exc_index* ix, const uint32_t* terms, uint32_t nt, uint32_t* out,
                              uint64_t cap, uint64_t* n_out)
{
    if (!c || !ix || (nt && !terms)) return set_error(EXC_ERR_ARG, "exc_decompress: null argument");
    uint64_t total;
    if (int rc = check_terms(ix, terms, nt, &total)) return rc;
    if (n_out) *n_out = total;
    if (total > cap) return set_error(EXC_ERR_CAPACITY, "output holds %llu, need %llu", (unsigned long long)cap,
                                      (unsigned long long)total);
    if (!total) return EXC_OK;
    if (!out) return set_error(EXC_ERR_ARG, "exc_decompress: null output");
    HIP_TRY(hipSetDevice(c->device));
    DevBuf dec;
    std::vector<uint64_t> off;
    int rc = decode_to_device(c, ix, std::vector<uint32_t>(terms, terms + nt), dec, off);
    if (!rc && hipMemcpy(out, dec.p, total * 4, hipMemcpyDeviceToHost) != hipSuccess)
        rc = set_error(EXC_ERR_DEVICE, "exc_decompress: copy to host failed");
    dev_free(c, dec);
    return rc;
}

extern "C" int exc_decompress_device(exc_ctx* c, exc_index* ix, const uint32_t* terms, uint32_t nt, uint32_t* d_out,
                                     uint64_t cap, uint64_t* n_out)
{
    if (!c || !ix || (nt && !terms)) return set_error(EXC_ERR_ARG, "exc_decompress_device: null argument");
    uint64_t total;
    if (int rc = check_terms(ix, terms, nt, &total)) return rc;
    if (n_out) *n_out = total;
    if (total > cap) return set_error(EXC_ERR_CAPACITY, "output holds %llu, need %llu", (unsigned long long)cap,
                                      (unsigned long long)total);
    if (!total) return EXC_OK;
    if (!d_out) return set_error(EXC_ERR_ARG, "exc_decompress_device: null output");
    HIP_TRY(hipSetDevice(c->device));
    DevBuf unused;
    std::vector<uint64_t> off;
    return decode_to_device(c, ix, std::vector<uint32_t>(terms, terms + nt), unused, off, d_out);
}

// ---------------------------------------------------------------- learned models

extern "C" int exc_index_train_models(exc_ctx* c, exc_index* ix, int kind, uint32_t min_len, uint32_t epochs,
                                      uint32_t samples, float lr, uint64_t seed, exc_model_stats* st)
{
    if (!c || !ix) return set_error(EXC_ERR_ARG, "exc_index_train_models: null argument");
    if (kind != EXC_MODEL_PLA && kind != EXC_MODEL_MLP && kind != EXC_MODEL_PLA_Q)
        return set_error(EXC_ERR_ARG, "unknown model kind %d", kind);
    if (ix->codec == EXC_HS_SEGLRC) return set_error(EXC_ERR_ARG, "HS_SegLRC lists are searched by bucket, not by position");
    if (ix->codec == EXC_NEWPFD) return set_error(EXC_ERR_ARG, "NewPFD indexes are for decompression only");
    if (kind == EXC_MODEL_MLP && (epochs == 0 || samples == 0)) return set_error(EXC_ERR_ARG, "MLP needs epochs and samples");
    HIP_TRY(hipSetDevice(c->device));
    std::vector<uint32_t> terms;
    for (uint32_t t = 0; t < ix->nterms; ++t)
        if (ix->n[t] > min_len && ix->n[t] >= 2) terms.push_back(t);
    const uint32_t nm = (uint32_t)terms.size();
    const uint32_t P = model_stride((uint32_t)kind);  // floats per model slot
    ix->models.assign(nm, Model{});
    for (uint32_t m = 0; m < nm; ++m) {
        Model& M = ix->models[m];
        std::memset(&M, 0, sizeof M);
        M.kind = (uint32_t)kind;
        M.n = ix->n[terms[m]];
        M.param_off = m * P;
        M.term = terms[m];
    }
    int rc;
    DevBuf errb;
    if ((rc = dev_alloc(c, ix->d_models, std::max<size_t>(1, nm) * sizeof(Model))) ||
        (rc = dev_alloc(c, ix->d_params, std::max<size_t>(1, (size_t)nm * P) * 4)) ||
        (rc = dev_alloc(c, errb, std::max<size_t>(1, nm) * 4)))
        return rc;
    if (nm) {
        HIP_TRY(hipMemcpy(ix->d_models.p, ix->models.data(), nm * sizeof(Model), hipMemcpyHostToDevice));
        HIP_TRY(hipMemset(errb.p, 0, nm * 4));
    }
    (void)hipEventRecord(c->ev[0], c->stream);
    const uint64_t kBatch = 64ull << 20;  // decoded elements per training batch
    for (uint32_t m0 = 0; m0 < nm;) {
        uint32_t m1 = m0;
        uint64_t tot = 0;
        while (m1 < nm && (m1 == m0 || tot + ix->n[terms[m1]] <= kBatch)) tot += ix->n[terms[m1++]];
        std::vector<uint32_t> bt(terms.begin() + m0, terms.begin() + m1);
        DevBuf dec, doff;
        std::vector<uint64_t> off;
        if ((rc = decode_to_device(c, ix, bt, dec, off)) || (rc = dev_alloc(c, doff, off.size() * 8))) {
            dev_free(c, dec);
            dev_free(c, errb);
            return rc;
        }
        hipError_t e = hipMemcpy(doff.p, off.data(), off.size() * 8, hipMemcpyHostToDevice);
        TrainArgs a;
        std::memset(&a, 0, sizeof a);
        a.dec = dptr<uint32_t>(dec);
        a.dec_off = dptr<uint64_t>(doff);
        a.models = dptr<Model>(ix->d_models) + m0;
        a.params = dptr<float>(ix->d_params);
        a.err_bits = dptr<uint32_t>(errb) + m0;
        a.nmodels = m1 - m0;
        a.epochs = epochs;
        a.samples = samples;
        a.lr = lr;
        a.seed = seed;
        if (e == hipSuccess)
            e = kind == EXC_MODEL_PLA     ? launch_pla_fit(a, c->stream)
                : kind == EXC_MODEL_PLA_Q ? launch_plaq_fit(a, c->stream)
                                          : launch_mlp_train(a, c->stream);
        if (e == hipSuccess) e = launch_model_err(a, tot, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        dev_free(c, dec);
        dev_free(c, doff);
        if (e != hipSuccess) {
            dev_free(c, errb);
            return set_error(EXC_ERR_DEVICE, "model training: %s", hipGetErrorString(e));
        }
        m0 = m1;
    }
    (void)hipEventRecord(c->ev[3], c->stream);
    (void)hipStreamSynchronize(c->stream);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[3]);
    std::vector<uint32_t> eb(nm);
    if (nm) {
        HIP_TRY(hipMemcpy(ix->models.data(), ix->d_models.p, nm * sizeof(Model), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(eb.data(), errb.p, nm * 4, hipMemcpyDeviceToHost));
    }
    dev_free(c, errb);
    double sum = 0;
    float mx = 0;
    for (uint32_t m = 0; m < nm; ++m) {
        float e;
        std::memcpy(&e, &eb[m], 4);
        ix->models[m].err = e;
        sum += e;
        mx = std::max(mx, e);
    }
    for (auto& R : ix->recs) {
        R.model = kNoModel;
        R.mkind = 0;
        R.merr = R.mx0 = R.mxs = 0.0f;
    }
    for (uint32_t m = 0; m < nm; ++m) {
        TermRec& R = ix->recs[terms[m]];
        const Model& M = ix->models[m];
        R.model = m;
        R.mkind = M.kind;
        R.merr = M.err;
        R.mx0 = M.x0;
        R.mxs = M.xs;
    }
    ++ix->model_gen;  // batch copies of the records are refreshed before their next launch
    if (nm) HIP_TRY(hipMemcpy(ix->d_models.p, ix->models.data(), nm * sizeof(Model), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ix->d_terms, ix->recs.data(), ix->recs.size() * sizeof(TermRec), hipMemcpyHostToDevice));
    ix->dix.models = dptr<Model>(ix->d_models);
    ix->dix.mparams = dptr<float>(ix->d_params);
    if (c->ix == ix) drop_graph(c);
    if (st) {
        st->nmodels = nm;
        st->max_err = mx;
        st->mean_err = nm ? (float)(sum / nm) : 0.0f;
        st->train_ms = ms;
    }
    return EXC_OK;
}

extern "C" int exc_index_model_info(exc_ctx* c, exc_index* ix, uint32_t term, exc_model_host* out)
{
    if (!c || !ix || !out || term >= ix->nterms) return set_error(EXC_ERR_ARG, "exc_index_model_info: bad argument");
    const uint32_t m = ix->recs[term].model;
    if (m == kNoModel || m >= ix->models.size()) return set_error(EXC_ERR_ARG, "term %u has no model", term);
    const Model& M = ix->models[m];
    std::memset(out, 0, sizeof *out);
    out->kind = (int32_t)M.kind;
    out->n = M.n;
    out->err = M.err;
    out->x0 = M.x0;
    out->xs = M.xs;
    out->nparams = model_nparams(M.kind);
    HIP_TRY(hipMemcpy(out->params, dptr<float>(ix->d_params) + M.param_off, out->nparams * 4, hipMemcpyDeviceToHost));
    return EXC_OK;
}

extern "C" int exc_mlp_gradient(exc_ctx* c, const float* params, const uint32_t* docs, uint32_t n, uint32_t samples,
                                float* grad, float* loss)
{
    if (!c || !params || !docs || !grad || n < 2 || samples == 0) return set_error(EXC_ERR_ARG, "exc_mlp_gradient: bad argument");
    HIP_TRY(hipSetDevice(c->device));
    DevBuf dl, doff, dm, dp, dg;
    int rc;
    if ((rc = dev_alloc(c, dl, (size_t)n * 4)) || (rc = dev_alloc(c, doff, 8)) || (rc = dev_alloc(c, dm, sizeof(Model))) ||
        (rc = dev_alloc(c, dp, 128 * 4)) || (rc = dev_alloc(c, dg, 128 * 4))) {
        for (DevBuf* b : {&dl, &doff, &dm, &dp, &dg}) dev_free(c, *b);
        return rc;
    }
    Model M;
    std::memset(&M, 0, sizeof M);
    M.kind = kModelMlp;
    M.n = n;
    const uint64_t zero = 0;
    hipError_t e = hipMemcpy(dl.p, docs, (size_t)n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(doff.p, &zero, 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dm.p, &M, sizeof M, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dp.p, params, kMlpParams * 4, hipMemcpyHostToDevice);
    TrainArgs a;
    std::memset(&a, 0, sizeof a);
    a.dec = dptr<uint32_t>(dl);
    a.dec_off = dptr<uint64_t>(doff);
    a.models = dptr<Model>(dm);
    a.params = dptr<float>(dp);
    a.nmodels = 1;
    a.epochs = 0;
    a.samples = samples;
    a.grad_out = dptr<float>(dg);
    if (e == hipSuccess) e = launch_mlp_train(a, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    float g[kMlpParams + 1];
    if (e == hipSuccess) e = hipMemcpy(g, dg.p, sizeof g, hipMemcpyDeviceToHost);
    for (DevBuf* b : {&dl, &doff, &dm, &dp, &dg}) dev_free(c, *b);
    if (e != hipSuccess) return set_error(EXC_ERR_DEVICE, "exc_mlp_gradient: %s", hipGetErrorString(e));
    std::memcpy(grad, g, kMlpParams * 4);
    if (loss) *loss = g[kMlpParams];
    return EXC_OK;
}

// ---------------------------------------------------------------- device index builder

namespace {

// Device buffers that live only during a build.
struct TempBufs {
    exc_ctx* c;
    std::vector<DevBuf> bufs;
    explicit TempBufs(exc_ctx* ctx) : c(ctx) {}
    ~TempBufs()
    {
        (void)hipStreamSynchronize(c->stream);
        for (auto& b : bufs) dev_free(c, b);
    }
    template <class T>
    int get(size_t count, T** p)
    {
        DevBuf b;
        if (int rc = dev_alloc(c, b, std::max<size_t>(1, count) * sizeof(T))) return rc;
        bufs.push_back(b);
        *p = static_cast<T*>(b.p);
        return EXC_OK;
    }
};

// An index array of `bytes` followed by kPadWords zero words; nothing copied in.
int alloc_padded(exc_index* ix, size_t bytes, void** out)
{
    DevBuf b;
    if (int rc = dev_alloc(ix->ctx, b, bytes + kPadWords * 4)) return rc;
    ix->bufs.push_back(b);
    ix->bytes += b.bytes;
    HIP_TRY(hipMemsetAsync(static_cast<char*>(b.p) + bytes, 0, kPadWords * 4, ix->ctx->stream));
    *out = b.p;
    return EXC_OK;
}

// Synthetic lists: items of ~4,096 expected docIDs per term, counted on the
// device; fills the host item table and the lists' offsets.
struct SynthPlan {
    std::vector<uint32_t> nchunks, item_term, item_chunk;
    std::vector<uint64_t> item_off, list_off;
    uint32_t *d_df = nullptr, *d_nch = nullptr, *d_it = nullptr, *d_ic = nullptr;
    uint64_t* d_ioff = nullptr;
};

int synth_plan(exc_ctx* c, TempBufs& tmp, uint64_t seed, uint32_t U, const uint32_t* df, uint32_t nterms,
               SynthPlan& P, SynthArgs& a)
{
    if (U == 0 || U > 0x80000000u) return set_error(EXC_ERR_ARG, "universe must be in [1, 2^31]");
    P.nchunks.resize(nterms);
    for (uint32_t t = 0; t < nterms; ++t) {
        if (df[t] > U) return set_error(EXC_ERR_ARG, "term %u: df %u exceeds the universe", t, df[t]);
        const uint32_t nc = std::max<uint32_t>(1, (df[t] + 4095) / 4096);
        P.nchunks[t] = nc;
        for (uint32_t k = 0; k < nc; ++k) {
            P.item_term.push_back(t);
            P.item_chunk.push_back(k);
        }
    }
    const uint64_t ni = P.item_term.size();
    uint32_t* d_cnt;
    int rc;
    if ((rc = tmp.get(nterms, &P.d_df)) || (rc = tmp.get(nterms, &P.d_nch)) || (rc = tmp.get(ni, &P.d_it)) ||
        (rc = tmp.get(ni, &P.d_ic)) || (rc = tmp.get(ni + 1, &P.d_ioff)) || (rc = tmp.get(ni, &d_cnt)))
        return rc;
    HIP_TRY(hipMemcpyAsync(P.d_df, df, nterms * 4ull, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(P.d_nch, P.nchunks.data(), nterms * 4ull, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(P.d_it, P.item_term.data(), ni * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(P.d_ic, P.item_chunk.data(), ni * 4, hipMemcpyHostToDevice, c->stream));
    std::memset(&a, 0, sizeof a);
    a.seed = seed;
    a.universe = U;
    a.df = P.d_df;
    a.nchunks = P.d_nch;
    a.item_term = P.d_it;
    a.item_chunk = P.d_ic;
    a.item_off = P.d_ioff;
    a.counts = d_cnt;
    a.item0 = 0;
    a.item1 = ni;
    HIP_TRY(launch_synth(a, false, c->stream));
    std::vector<uint32_t> cnt(ni);
    HIP_TRY(hipMemcpyAsync(cnt.data(), d_cnt, ni * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    P.item_off.assign(ni + 1, 0);
    for (uint64_t i = 0; i < ni; ++i) P.item_off[i + 1] = P.item_off[i] + cnt[i];
    P.list_off.assign(nterms + 1, 0);
    uint64_t i = 0;
    for (uint32_t t = 0; t < nterms; ++t) {
        P.list_off[t] = P.item_off[i];
        i += P.nchunks[t];
        if (P.item_off[i] - P.list_off[t] > 0xffffffffull) return set_error(EXC_ERR_CAPACITY, "term %u: list too long", t);
    }
    P.list_off[nterms] = P.item_off[ni];
    HIP_TRY(hipMemcpyAsync(P.d_ioff, P.item_off.data(), (ni + 1) * 8, hipMemcpyHostToDevice, c->stream));
    return EXC_OK;
}

// Write the docIDs of terms [t0, t1) into docs (docs[0] = list t0's first).
int synth_write(exc_ctx* c, const SynthPlan& P, SynthArgs a, uint32_t t0, uint32_t t1, uint32_t* docs)
{
    uint64_t i0 = 0;
    for (uint32_t t = 0; t < t0; ++t) i0 += P.nchunks[t];
    uint64_t i1 = i0;
    for (uint32_t t = t0; t < t1; ++t) i1 += P.nchunks[t];
    a.docs = docs;
    a.doc_base = P.list_off[t0];
    a.item0 = i0;
    a.item1 = i1;
    HIP_TRY(launch_synth(a, true, c->stream));
    return EXC_OK;
}

// SegLRC, LRC or ParaPFD encoding on the device, block for block the host
// builder's bytes.  docs_for(t0, t1, &ptr) makes the docIDs of lists [t0, t1)
// available in device memory (ptr[0] = list t0's first); it is called twice
// per batch (fits and block sizes, then packing).
template <class DocsFor>
int build_index_on_device(exc_ctx* c, int codec, uint32_t seg, float frac, uint32_t nterms,
                          const std::vector<uint64_t>& list_off,
                          const std::vector<std::pair<uint32_t, uint32_t>>& batches, DocsFor docs_for, exc_index** out)
{
    if (seg != 128 && seg != 256) return set_error(EXC_ERR_ARG, "segment size must be 128 or 256");
    if (codec != EXC_SEGLRC && codec != EXC_LRC && codec != EXC_PARAPFD)
        return set_error(EXC_ERR_ARG, "the device builder encodes SegLRC, LRC and ParaPFD");
    if (codec != EXC_PARAPFD) frac = 0.0f;  // the LRC-family probes read blocks without exceptions
    if (!(frac >= 0.0f && frac <= 1.0f)) return set_error(EXC_ERR_ARG, "FRAC must be in [0, 1]");
    std::vector<uint32_t> seg_num(nterms + 1, 0);
    uint64_t S = 0;
    for (uint32_t t = 0; t < nterms; ++t) {
        seg_num[t] = (uint32_t)S;
        S += (list_off[t + 1] - list_off[t] + seg - 1) / seg;
        if (S > 0xfffffffeull) return set_error(EXC_ERR_CAPACITY, "more than 2^32 segments");
    }
    seg_num[nterms] = (uint32_t)S;
    TempBufs tmp(c);
    auto* ix = new exc_index();
    ix->ctx = c;
    ix->codec = codec;
    ix->seg = seg;
    ix->nterms = nterms;
    ix->nseg = S;
    int rc = EXC_OK;
    auto fail = [&](int r) {
        exc_index_free(ix);
        return r;
    };
    uint64_t* d_loff;
    uint32_t *d_snum, *d_nw;
    uint64_t* d_goff;
    uint32_t* d_err;
    void *p_rec, *p_head, *p_terms, *p_oct = nullptr;
    if ((rc = tmp.get(nterms + 1, &d_loff)) || (rc = tmp.get(nterms + 1, &d_snum)) || (rc = tmp.get(S + 1, &d_nw)) ||
        (rc = tmp.get(S + 1, &d_goff)) || (rc = tmp.get(1, &d_err)) || (rc = alloc_padded(ix, S * sizeof(SegRec), &p_rec)) ||
        (rc = alloc_padded(ix, S * 4, &p_head)) || (rc = alloc_padded(ix, (size_t)nterms * sizeof(TermRec), &p_terms)) ||
        (codec == EXC_PARAPFD && (rc = alloc_padded(ix, S * 16, &p_oct))))
        return fail(rc);
    hipError_t e = hipMemcpyAsync(d_loff, list_off.data(), (nterms + 1ull) * 8, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_snum, seg_num.data(), (nterms + 1ull) * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(d_nw + S, 0, 4, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(d_err, 0, 4, c->stream);
    if (e != hipSuccess) return fail(set_error(EXC_ERR_DEVICE, "device build: %s", hipGetErrorString(e)));
    EncodeArgs a;
    std::memset(&a, 0, sizeof a);
    a.list_off = d_loff;
    a.seg_num = d_snum;
    a.nterms = nterms;
    a.seg = seg;
    a.rec = static_cast<SegRec*>(p_rec);
    a.head = static_cast<uint32_t*>(p_head);
    a.nwords = d_nw;
    a.goff = d_goff;
    a.terms = static_cast<TermRec*>(p_terms);
    a.pfd_oct = static_cast<uint32_t*>(p_oct);
    a.codec = codec;
    a.frac = frac;
    a.err = d_err;
    // pass 1: fits and block sizes
    for (auto [t0, t1] : batches) {
        const uint32_t* docs;
        if ((rc = docs_for(t0, t1, &docs))) return fail(rc);
        a.docs = docs;
        a.doc_base = list_off[t0];
        a.g0 = seg_num[t0];
        a.g1 = seg_num[t1];
        a.t0 = t0;
        a.t1 = t1;
        if (codec == EXC_SEGLRC) {
            e = launch_seg_fit(a, c->stream);
        } else {
            e = codec == EXC_LRC ? launch_lrc_lists(a, c->stream) : hipSuccess;
            if (e == hipSuccess) e = launch_blk_fit(a, c->stream);
        }
        if (e != hipSuccess) break;
    }
    // block offsets, words, term records
    uint64_t W = 0;
    uint32_t berr = 0;
    if (e == hipSuccess) e = scan_u32_to_u64(d_nw, d_goff, S + 1, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&W, d_goff + S, 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&berr, d_err, 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(set_error(EXC_ERR_DEVICE, "device build: %s", hipGetErrorString(e)));
    if (berr) return fail(set_error(EXC_ERR_FORMAT, "device build: a block cannot be encoded (docIDs must be < 2^31)"));
    void* p_words;
    if ((rc = alloc_padded(ix, W * 4, &p_words))) return fail(rc);
    a.words = static_cast<uint32_t*>(p_words);
    if ((e = launch_term_recs(a, c->stream)) != hipSuccess)
        return fail(set_error(EXC_ERR_DEVICE, "device build: %s", hipGetErrorString(e)));
    // pass 2: packed blocks, list-level fits
    for (auto [t0, t1] : batches) {
        const uint32_t* docs;
        if ((rc = docs_for(t0, t1, &docs))) return fail(rc);
        a.docs = docs;
        a.doc_base = list_off[t0];
        a.g0 = seg_num[t0];
        a.g1 = seg_num[t1];
        a.t0 = t0;
        a.t1 = t1;
        if ((e = (codec == EXC_SEGLRC ? launch_seg_pack(a, c->stream) : launch_blk_pack(a, c->stream))) != hipSuccess)
            break;
    }
    if (e == hipSuccess && codec == EXC_PARAPFD) e = launch_heads_line(a, c->stream);
    ix->recs.resize(nterms);
    if (e == hipSuccess)
        e = hipMemcpyAsync(ix->recs.data(), p_terms, (size_t)nterms * sizeof(TermRec), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(set_error(EXC_ERR_DEVICE, "device build: %s", hipGetErrorString(e)));
    ix->nwords = W;
    ix->n.resize(nterms);
    for (uint32_t t = 0; t < nterms; ++t) ix->n[t] = ix->recs[t].n;
    // the device builder packs the lists back to back in term order
    ix->lwords.resize(nterms);
    for (uint32_t t = 0; t < nterms; ++t)
        ix->lwords[t] = (uint32_t)((t + 1 < nterms ? ix->recs[t + 1].woff : W) - ix->recs[t].woff);
    ix->dix.codec = codec;
    ix->dix.seg = seg;
    // k_blk_fit keeps each block's median, its other quarter values and the
    // quarters' midpoint values
    ix->dix.flags = codec == EXC_PARAPFD ? kIxMedians | kIxQuarters | kIxOctants : 0u;
    ix->dix.pfd_oct = a.pfd_oct;
    ix->dix.words = a.words;
    ix->dix.terms = a.terms;
    ix->d_terms = a.terms;
    ix->dix.rec = a.rec;
    ix->dix.seg_head = a.head;
    *out = ix;
    return EXC_OK;
}

// HS_SegLRC encoding on the device, bucket for bucket the host builder's bytes
// (hash tables, fits, bases, packed residuals): the lists' shifts and bucket
// counts first (they depend on each list's maximum), then the buckets, then
// after the word-offset scan the packing.  docs_for as for the other codecs,
// called three times per batch.
template <class DocsFor>
int build_hs_on_device(exc_ctx* c, uint32_t hs_target, uint32_t hs_shift, uint32_t nterms,
                       const std::vector<uint64_t>& list_off, const std::vector<std::pair<uint32_t, uint32_t>>& batches,
                       DocsFor docs_for, exc_index** out)
{
    if (hs_target == 0) hs_target = 256;
    if (hs_shift > 31) return set_error(EXC_ERR_ARG, "hash shift %u above 31", hs_shift);
    TempBufs tmp(c);
    uint64_t* d_loff;
    uint32_t *d_shift, *d_nb;
    int rc;
    if ((rc = tmp.get(nterms + 1, &d_loff)) || (rc = tmp.get(std::max<uint32_t>(nterms, 1), &d_shift)) ||
        (rc = tmp.get(std::max<uint32_t>(nterms, 1), &d_nb)))
        return rc;
    HIP_TRY(hipMemcpyAsync(d_loff, list_off.data(), (nterms + 1ull) * 8, hipMemcpyHostToDevice, c->stream));
    EncodeArgs a;
    std::memset(&a, 0, sizeof a);
    a.list_off = d_loff;
    a.nterms = nterms;
    a.seg = 128;
    a.codec = EXC_HS_SEGLRC;
    a.hs_shift = d_shift;
    a.hs_nb = d_nb;
    a.hs_target = hs_target;
    a.hs_fixed = hs_shift;
    // pass 1: shifts and bucket counts
    for (auto [t0, t1] : batches) {
        const uint32_t* docs;
        if ((rc = docs_for(t0, t1, &docs))) return rc;
        a.docs = docs;
        a.doc_base = list_off[t0];
        a.t0 = t0;
        a.t1 = t1;
        HIP_TRY(launch_hs_lists(a, c->stream));
    }
    std::vector<uint32_t> nb(nterms);
    if (nterms) HIP_TRY(hipMemcpyAsync(nb.data(), d_nb, nterms * 4ull, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    std::vector<uint32_t> seg_num(nterms + 1, 0);
    std::vector<uint64_t> hoff(nterms + 1, 0);
    for (uint32_t t = 0; t < nterms; ++t) {
        seg_num[t + 1] = seg_num[t] + nb[t];
        hoff[t + 1] = hoff[t] + nb[t] + 3ull;
        if ((uint64_t)seg_num[t] + nb[t] > 0xfffffffeull || hoff[t + 1] > 0xffffffffull)
            return set_error(EXC_ERR_CAPACITY, "more than 2^32 buckets or table words");
    }
    const uint64_t S = seg_num[nterms], H = hoff[nterms];
    auto* ix = new exc_index();
    ix->ctx = c;
    ix->codec = EXC_HS_SEGLRC;
    ix->seg = 128;
    ix->nterms = nterms;
    ix->nseg = S;
    auto fail = [&](int r) {
        exc_index_free(ix);
        return r;
    };
    uint32_t *d_snum, *d_nw, *p_hash;
    uint64_t *d_goff, *d_hoff;
    SegRec* p_rec;
    void* p_terms;
    // every bucket's record and the tables are temporaries: the index keeps
    // the compact layout made from them below
    if ((rc = tmp.get(nterms + 1, &d_snum)) || (rc = tmp.get(S + 1, &d_nw)) || (rc = tmp.get(S + 1, &d_goff)) ||
        (rc = tmp.get(nterms + 1, &d_hoff)) || (rc = tmp.get(S, &p_rec)) || (rc = tmp.get(H, &p_hash)) ||
        (rc = alloc_padded(ix, (size_t)nterms * sizeof(TermRec), &p_terms)))
        return fail(rc);
    hipError_t e = hipMemcpyAsync(d_snum, seg_num.data(), (nterms + 1ull) * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_hoff, hoff.data(), (nterms + 1ull) * 8, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(d_nw + S, 0, 4, c->stream);
    if (e != hipSuccess) return fail(set_error(EXC_ERR_DEVICE, "device build: %s", hipGetErrorString(e)));
    a.seg_num = d_snum;
    a.rec = p_rec;
    a.nwords = d_nw;
    a.goff = d_goff;
    a.terms = static_cast<TermRec*>(p_terms);
    a.hs_off = d_hoff;
    a.hash = p_hash;
    // pass 2: tables, fits, bases, block sizes
    for (auto [t0, t1] : batches) {
        const uint32_t* docs;
        if ((rc = docs_for(t0, t1, &docs))) return fail(rc);
        a.docs = docs;
        a.doc_base = list_off[t0];
        a.g0 = seg_num[t0];
        a.g1 = seg_num[t1];
        if ((e = launch_hs_fit(a, c->stream)) != hipSuccess) break;
    }
    uint64_t W = 0;
    if (e == hipSuccess) e = scan_u32_to_u64(d_nw, d_goff, S + 1, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&W, d_goff + S, 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(set_error(EXC_ERR_DEVICE, "device build: %s", hipGetErrorString(e)));
    void* p_words;
    if ((rc = alloc_padded(ix, W * 4, &p_words))) return fail(rc);
    a.words = static_cast<uint32_t*>(p_words);
    if ((e = launch_term_recs(a, c->stream)) != hipSuccess)
        return fail(set_error(EXC_ERR_DEVICE, "device build: %s", hipGetErrorString(e)));
    // pass 3: packed residuals
    for (auto [t0, t1] : batches) {
        const uint32_t* docs;
        if ((rc = docs_for(t0, t1, &docs))) return fail(rc);
        a.docs = docs;
        a.doc_base = list_off[t0];
        a.g0 = seg_num[t0];
        a.g1 = seg_num[t1];
        if ((e = launch_hs_pack(a, c->stream)) != hipSuccess) break;
    }
    // compaction: records of the non-empty buckets, occupancy tables of the
    // lists that have empty buckets
    HsCompactArgs k;
    std::memset(&k, 0, sizeof k);
    k.seg_num = d_snum;
    k.hs_off = d_hoff;
    k.hash = p_hash;
    k.rec = p_rec;
    k.nterms = nterms;
    k.nbuckets = S;
    uint32_t *d_flag, *d_count;
    uint64_t *d_cidx, *d_occoff, *d_grpoff;
    if ((rc = tmp.get(S + 1, &d_flag)) || (rc = tmp.get(S + 1, &d_cidx)) ||
        (rc = tmp.get(std::max<uint32_t>(nterms, 1), &d_count)) || (rc = tmp.get(std::max<uint32_t>(nterms, 1), &d_occoff)) ||
        (rc = tmp.get(nterms + 1, &d_grpoff)))
        return fail(rc);
    k.flag = d_flag;
    k.cidx = d_cidx;
    k.count = d_count;
    std::vector<uint32_t> count(nterms);
    uint64_t C = 0;
    if (e == hipSuccess) e = launch_hs_compact_flags(k, c->stream);
    if (e == hipSuccess) e = scan_u32_to_u64(d_flag, d_cidx, S + 1, c->stream);
    if (e == hipSuccess) e = launch_hs_compact_counts(k, c->stream);
    if (e == hipSuccess && nterms) e = hipMemcpyAsync(count.data(), d_count, nterms * 4ull, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&C, d_cidx + S, 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(set_error(EXC_ERR_DEVICE, "device build: %s", hipGetErrorString(e)));
    std::vector<uint64_t> occ_off(std::max<uint32_t>(nterms, 1), ~0ull), grp_off(nterms + 1, 0);
    uint64_t O = 0;
    for (uint32_t t = 0; t < nterms; ++t) {
        const bool dense = count[t] == nb[t];
        grp_off[t + 1] = grp_off[t] + (dense ? 0u : (nb[t] + 31) / 32);
        if (!dense) {
            occ_off[t] = O;
            O += hs_occ_words(nb[t]);
        }
    }
    if (C > 0xffffffffull || O > 0xfffffffeull)
        return fail(set_error(EXC_ERR_CAPACITY, "more than 2^32 records or table words"));
    void *p_crec, *p_chead, *p_occ;
    if ((rc = alloc_padded(ix, C * sizeof(SegRec), &p_crec)) || (rc = alloc_padded(ix, C * 4, &p_chead)) ||
        (rc = alloc_padded(ix, O * 4, &p_occ)))
        return fail(rc);
    k.occ_off = d_occoff;
    k.grp_off = d_grpoff;
    k.ngroups = grp_off[nterms];
    k.crec = static_cast<SegRec*>(p_crec);
    k.chead = static_cast<uint32_t*>(p_chead);
    k.occ = static_cast<uint32_t*>(p_occ);
    k.terms = a.terms;
    e = hipMemcpyAsync(d_occoff, occ_off.data(), occ_off.size() * 8, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_grpoff, grp_off.data(), grp_off.size() * 8, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = launch_hs_compact(k, c->stream);
    ix->recs.resize(nterms);
    if (e == hipSuccess && nterms)
        e = hipMemcpyAsync(ix->recs.data(), p_terms, (size_t)nterms * sizeof(TermRec), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(set_error(EXC_ERR_DEVICE, "device build: %s", hipGetErrorString(e)));
    ix->nwords = W;
    ix->nseg = C;
    ix->hash_words = O;
    ix->n.resize(nterms);
    for (uint32_t t = 0; t < nterms; ++t) ix->n[t] = ix->recs[t].n;
    ix->dix.codec = EXC_HS_SEGLRC;
    ix->dix.seg = 128;
    ix->dix.words = a.words;
    ix->dix.terms = a.terms;
    ix->d_terms = a.terms;
    ix->dix.rec = k.crec;
    ix->dix.seg_head = k.chead;
    ix->dix.hash = k.occ;
    *out = ix;
    return EXC_OK;
}

}  // namespace

extern "C" int exc_synth_lists_device(exc_ctx* c, uint64_t seed, uint32_t universe, const uint32_t* df, uint32_t nterms,
                                      uint64_t* list_off, uint32_t* d_docs, uint64_t cap)
{
    if (!c || (nterms && !df) || !list_off) return set_error(EXC_ERR_ARG, "exc_synth_lists_device: bad argument");
    HIP_TRY(hipSetDevice(c->device));
    TempBufs tmp(c);
    SynthPlan P;
    SynthArgs a;
    if (int rc = synth_plan(c, tmp, seed, universe, df, nterms, P, a)) return rc;
    std::memcpy(list_off, P.list_off.data(), (nterms + 1ull) * 8);
    if (!d_docs) return EXC_OK;
    if (P.list_off[nterms] > cap)
        return set_error(EXC_ERR_CAPACITY, "output holds %llu docIDs, need %llu", (unsigned long long)cap,
                         (unsigned long long)P.list_off[nterms]);
    if (int rc = synth_write(c, P, a, 0, nterms, d_docs)) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return EXC_OK;
}

extern "C" int exc_build_index_device(exc_ctx* c, int codec, uint32_t seg, float frac, uint32_t hs_target,
                                      uint32_t hs_shift, const uint32_t* d_docs, const uint64_t* list_off,
                                      uint32_t nterms, exc_index** out)
{
    if (!c || !list_off || !out || (list_off[nterms] && !d_docs))
        return set_error(EXC_ERR_ARG, "exc_build_index_device: bad argument");
    for (uint32_t t = 0; t < nterms; ++t)
        if (list_off[t + 1] < list_off[t] || list_off[t + 1] - list_off[t] > 0xffffffffull)
            return set_error(EXC_ERR_ARG, "exc_build_index_device: bad list offsets");
    HIP_TRY(hipSetDevice(c->device));
    std::vector<uint64_t> lo(list_off, list_off + nterms + 1);
    std::vector<std::pair<uint32_t, uint32_t>> batches{{0u, nterms}};
    const auto docs_for = [&](uint32_t t0, uint32_t, const uint32_t** p) {
        *p = d_docs + lo[t0];
        return EXC_OK;
    };
    if (codec == EXC_HS_SEGLRC) return build_hs_on_device(c, hs_target, hs_shift, nterms, lo, batches, docs_for, out);
    return build_index_on_device(c, codec, seg, frac, nterms, lo, batches, docs_for, out);
}

extern "C" int exc_synth_index_device(exc_ctx* c, int codec, uint32_t seg, float frac, uint32_t hs_target,
                                      uint32_t hs_shift, uint64_t seed, uint32_t universe, const uint32_t* df,
                                      uint32_t nterms, uint64_t batch_postings, uint64_t* list_off_out,
                                      exc_index** out)
{
    if (!c || (nterms && !df) || !out) return set_error(EXC_ERR_ARG, "exc_synth_index_device: bad argument");
    HIP_TRY(hipSetDevice(c->device));
    TempBufs tmp(c);
    SynthPlan P;
    SynthArgs a;
    if (int rc = synth_plan(c, tmp, seed, universe, df, nterms, P, a)) return rc;
    if (list_off_out) std::memcpy(list_off_out, P.list_off.data(), (nterms + 1ull) * 8);
    if (batch_postings == 0) batch_postings = 1ull << 30;
    std::vector<std::pair<uint32_t, uint32_t>> batches;
    uint64_t maxb = 1;
    for (uint32_t t0 = 0; t0 < nterms;) {
        uint32_t t1 = t0 + 1;
        while (t1 < nterms && P.list_off[t1 + 1] - P.list_off[t0] <= batch_postings) ++t1;
        batches.emplace_back(t0, t1);
        maxb = std::max<uint64_t>(maxb, P.list_off[t1] - P.list_off[t0]);
        t0 = t1;
    }
    uint32_t* d_docs;
    if (int rc = tmp.get(maxb, &d_docs)) return rc;
    const auto docs_for = [&](uint32_t t0, uint32_t t1, const uint32_t** p) {
        *p = d_docs;
        return synth_write(c, P, a, t0, t1, d_docs);
    };
    if (codec == EXC_HS_SEGLRC)
        return build_hs_on_device(c, hs_target, hs_shift, nterms, P.list_off, batches, docs_for, out);
    return build_index_on_device(c, codec, seg, frac, nterms, P.list_off, batches, docs_for, out);
}

extern "C" int exc_index_export(exc_ctx* c, exc_index* ix, exc_built** out)
{
    if (!c || !ix || !out) return set_error(EXC_ERR_ARG, "exc_index_export: bad argument");
    if (ix->codec != EXC_SEGLRC && ix->codec != EXC_LRC && ix->codec != EXC_PARAPFD && ix->codec != EXC_HS_SEGLRC)
        return set_error(EXC_ERR_ARG, "exc_index_export: SegLRC, LRC, ParaPFD and HS_SegLRC indexes");
    HIP_TRY(hipSetDevice(c->device));
    const bool hs = ix->codec == EXC_HS_SEGLRC;
    std::vector<SegRec> rec(ix->nseg);
    std::vector<uint32_t> head(hs ? 0 : ix->nseg), words(ix->nwords), occ(hs ? ix->hash_words : 0);
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (ix->nseg) {
        HIP_TRY(hipMemcpy(rec.data(), ix->dix.rec, ix->nseg * sizeof(SegRec), hipMemcpyDeviceToHost));
        if (!hs) HIP_TRY(hipMemcpy(head.data(), ix->dix.seg_head, ix->nseg * 4, hipMemcpyDeviceToHost));
    }
    if (ix->nwords) HIP_TRY(hipMemcpy(words.data(), ix->dix.words, ix->nwords * 4, hipMemcpyDeviceToHost));
    if (!occ.empty()) HIP_TRY(hipMemcpy(occ.data(), ix->dix.hash, occ.size() * 4, hipMemcpyDeviceToHost));
    if (!hs)
        return built_from_device(ix->codec, ix->seg, ix->nterms, ix->recs.data(), ix->nseg, rec.data(), head.data(),
                                 words.data(), ix->nwords, nullptr, 0, out);
    std::vector<TermRec> fterms;
    std::vector<SegRec> frec;
    std::vector<uint32_t> table;
    hs_expand(ix->recs, rec, occ, words, fterms, frec, table);
    return built_from_device(ix->codec, ix->seg, ix->nterms, fterms.data(), frec.size(), frec.data(), nullptr,
                             words.data(), ix->nwords, table.data(), table.size(), out);
}

// ---------------------------------------------------------------- autotuner

namespace {

// Best probe time over `reps` runs of the planned batch (after one warm run).
int time_probe(exc_ctx* c, uint32_t reps, float* best)
{
    int rc = upload_plan(c);
    if (rc) return rc;
    if ((rc = invoke(c))) return rc;
    *best = 1e30f;
    for (uint32_t r = 0; r < reps; ++r) {
        if ((rc = invoke(c))) return rc;
        *best = std::min(*best, c->probe_ms);
    }
    return EXC_OK;
}

}  // namespace

extern "C" int exc_autotune(exc_ctx* c, exc_index* ix, const uint32_t* qterms, const uint64_t* qoff, uint32_t nq,
                            uint32_t reps, exc_tune_table* out)
{
    if (!c || !ix || !qoff || (nq && !qterms) || !out) return set_error(EXC_ERR_ARG, "exc_autotune: bad argument");
    if (ix->codec == EXC_NEWPFD) return set_error(EXC_ERR_ARG, "NewPFD indexes are for decompression only");
    reps = std::max<uint32_t>(1, std::min<uint32_t>(reps, 50));
    // classify the sample
    std::vector<std::vector<uint32_t>> cls(kRatioClasses);
    std::vector<uint32_t> tmp;
    for (uint32_t q = 0; q < nq; ++q) {
        if (qoff[q + 1] < qoff[q]) return set_error(EXC_ERR_ARG, "query %u: bad offsets", q);
        tmp.assign(qterms + qoff[q], qterms + qoff[q + 1]);
        for (uint32_t t : tmp)
            if (t >= ix->nterms) return set_error(EXC_ERR_ARG, "query %u: term %u out of range", q, t);
        std::stable_sort(tmp.begin(), tmp.end(), [&](uint32_t a, uint32_t b) { return ix->n[a] < ix->n[b]; });
        cls[ratio_class(ix, tmp.data(), (uint32_t)tmp.size())].push_back(q);
    }
    const bool had_graph = c->graph, had_timing = c->timing;
    c->graph = 0;
    c->timing = 1;
    drop_graph(c);
    exc_tune_table T;
    std::memset(&T, 0, sizeof T);
    // candidates: cooperative-run limits per codec x search modes
    std::vector<uint32_t> coops;
    if (ix->codec == EXC_PARAPFD) coops = {16, 32, 64};
    else if (ix->codec == EXC_HS_SEGLRC || ix->codec == EXC_RAW || ix->codec == EXC_RAW_HS) coops = {0};
    else coops = {0, 2, 4, 8, 16, 32, 64};
    std::vector<uint32_t> searches = {SEARCH_BINARY, SEARCH_FIT};
    if (ix->codec == EXC_RAW) searches.push_back(SEARCH_INTERP);
    if (!ix->models.empty()) searches.push_back(SEARCH_LEARNED);
    int rc = EXC_OK;
    for (int k = 0; k < kRatioClasses && !rc; ++k) {
        c->tuned = false;
        const uint32_t dk = knobs_for(c, ix->codec, (uint32_t)k);
        T.coop_max[k] = knob_coop(dk);
        T.search[k] = knob_search(dk);
        T.queries[k] = (uint32_t)cls[k].size();
        if (cls[k].empty() || k == kRatioClasses - 1) continue;  // single-term queries probe nothing
        // sub-batch of the class, at most 4096 queries spread over the sample
        std::vector<uint32_t> sq;
        std::vector<uint64_t> so{0};
        const size_t m = cls[k].size(), take = std::min<size_t>(m, 4096);
        for (size_t i = 0; i < take; ++i) {
            const uint32_t q = cls[k][i * m / take];
            sq.insert(sq.end(), qterms + qoff[q], qterms + qoff[q + 1]);
            so.push_back(sq.size());
        }
        std::vector<uint32_t> cands;
        for (uint32_t s : searches)
            for (uint32_t cp : coops) cands.push_back(make_knobs(cp, s));
        if (std::find(cands.begin(), cands.end(), dk) == cands.end()) cands.push_back(dk);
        float best = 1e30f;
        for (uint32_t kn : cands) {
            c->tuned = true;
            for (int j = 0; j < kRatioClasses; ++j) {
                c->tune.coop_max[j] = knob_coop(kn);
                c->tune.search[j] = knob_search(kn);
            }
            float ms = 0;
            if ((rc = plan_queries(c, ix, sq.data(), so.data(), (uint32_t)take)) ||
                (rc = time_probe(c, reps, &ms)))
                break;
            if (kn == dk) T.ms_default[k] = ms;
            if (ms < best) {
                best = ms;
                T.coop_max[k] = knob_coop(kn);
                T.search[k] = knob_search(kn);
            }
        }
        T.ms[k] = best;
    }
    // launch mode over the whole sample with the chosen knobs
    const int had_persistent = c->persistent;
    if (!rc) {
        c->tuned = true;
        c->tune = T;
        for (int p = 0; p < 2 && !rc; ++p) {
            c->persistent = p;
            if (!(rc = plan_queries(c, ix, qterms, qoff, nq))) rc = time_probe(c, reps, &T.batch_ms[p]);
        }
        T.persistent = T.batch_ms[1] < T.batch_ms[0] ? 1u : 0u;
    }
    c->graph = had_graph;
    c->timing = had_timing;
    c->state = 0;  // the last plan was a tuning batch
    if (rc) {
        c->tuned = false;
        c->persistent = had_persistent;
        return rc;
    }
    c->persistent = (int)T.persistent;
    c->tune = T;
    c->tuned = true;
    ++c->opt_gen;
    *out = T;
    return EXC_OK;
}

extern "C" int exc_ctx_set_tuning(exc_ctx* c, const exc_tune_table* t)
{
    if (!c) return set_error(EXC_ERR_ARG, "exc_ctx_set_tuning: null ctx");
    if (t) {
        for (int k = 0; k < kRatioClasses; ++k)
            if (t->coop_max[k] > 64 || t->search[k] > 3)
                return set_error(EXC_ERR_ARG, "tuning table class %d: coop_max <= 64 and search <= 3", k);
        c->tune = *t;
        c->persistent = t->persistent != 0;
    }
    c->tuned = t != nullptr;
    ++c->opt_gen;
    drop_graph(c);
    return EXC_OK;
}

// ---------------------------------------------------------------- router

// One context per candidate: its own stream and batch buffers, so the
// candidates' sub-batches are queued side by side and run concurrently.  The
// router's own stream forks them (each candidate stream waits on a marker of
// it) and joins them (it waits on a marker of each), which also makes the
// whole run one capturable unit.
struct exc_router {
    exc_ctx* parent = nullptr;
    std::vector<exc_index*> cands;
    std::vector<exc_ctx*> kids;
    uint32_t route[kRatioClasses] = {};
    hipStream_t stream = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};  // around the run, on the router stream
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    int graph = 0;
    hipGraphExec_t gexec = nullptr;
    std::vector<uint64_t> graph_model_gen;  // each candidate's model generation the graph captured
    // plan of the batch
    uint32_t nq = 0;
    int state = 0;  // 0 none, 2 planned and on the device, 3 run
    std::vector<uint8_t> active;  // candidate has queries in the batch
    std::vector<uint32_t> qcand, qidx;  // candidate of each query, its number in that sub-batch
    float ms = 0;
};

namespace {

void router_drop_graph(exc_router* r)
{
    if (r->gexec) (void)hipGraphExecDestroy(r->gexec);
    r->gexec = nullptr;
}

int router_plan(exc_router* r, const uint32_t* qterms, const uint64_t* qoff, uint32_t nq)
{
    const size_t n = r->cands.size();
    const exc_index* ix0 = r->cands[0];
    std::vector<std::vector<uint32_t>> st(n);
    std::vector<std::vector<uint64_t>> so(n, std::vector<uint64_t>{0});
    r->state = 0;
    r->qcand.resize(nq);
    r->qidx.resize(nq);
    std::vector<uint32_t> tmp;
    for (uint32_t q = 0; q < nq; ++q) {
        if (qoff[q + 1] < qoff[q]) return set_error(EXC_ERR_ARG, "query %u: bad offsets", q);
        tmp.assign(qterms + qoff[q], qterms + qoff[q + 1]);
        for (uint32_t t : tmp)
            if (t >= ix0->nterms) return set_error(EXC_ERR_ARG, "query %u: term %u out of range", q, t);
        std::stable_sort(tmp.begin(), tmp.end(), [&](uint32_t a, uint32_t b) { return ix0->n[a] < ix0->n[b]; });
        const uint32_t i = r->route[ratio_class(ix0, tmp.data(), (uint32_t)tmp.size())];
        r->qcand[q] = i;
        r->qidx[q] = (uint32_t)(so[i].size() - 1);
        st[i].insert(st[i].end(), qterms + qoff[q], qterms + qoff[q + 1]);
        so[i].push_back(st[i].size());
    }
    router_drop_graph(r);
    for (size_t i = 0; i < n; ++i) {
        r->active[i] = so[i].size() > 1;
        if (!r->active[i]) continue;
        int rc = plan_queries(r->kids[i], r->cands[i], st[i].data(), so[i].data(), (uint32_t)(so[i].size() - 1));
        if (!rc) rc = upload_plan(r->kids[i]);
        if (rc) return rc;
    }
    r->nq = nq;
    r->state = 2;
    return EXC_OK;
}

// Queue every active candidate's sub-batch between the fork and the join.
int router_enqueue(exc_router* r, hipEvent_t fork)
{
    for (size_t i = 0; i < r->kids.size(); ++i) {
        if (!r->active[i]) continue;
        exc_ctx* k = r->kids[i];
        HIP_TRY(hipStreamWaitEvent(k->stream, fork, 0));
        if (int rc = enqueue_invoke(k, false)) return rc;
        HIP_TRY(hipEventRecord(r->join[i], k->stream));
        HIP_TRY(hipStreamWaitEvent(r->stream, r->join[i], 0));
    }
    return EXC_OK;
}

int router_run(exc_router* r)
{
    if (r->state < 2) return set_error(EXC_ERR_STATE, "exc_router_invoke before exc_router_plan");
    HIP_TRY(hipSetDevice(r->parent->device));
    if (std::find(r->active.begin(), r->active.end(), 1) == r->active.end()) {  // an empty batch
        r->ms = 0;
        r->state = 3;
        return EXC_OK;
    }
    for (size_t i = 0; i < r->kids.size(); ++i) {
        if (!r->active[i]) continue;
        if (int rc = refresh_knobs(r->kids[i])) return rc;
        if (r->gexec && r->graph_model_gen[i] != r->cands[i]->model_gen) router_drop_graph(r);
    }
    if (r->graph) {
        if (!r->gexec) {
            hipGraph_t g = nullptr;
            HIP_TRY(hipStreamBeginCapture(r->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = hipEventRecord(r->fork, r->stream);
            int rc = e == hipSuccess ? router_enqueue(r, r->fork)
                                     : set_error(EXC_ERR_DEVICE, "hipEventRecord: %s", hipGetErrorString(e));
            e = hipStreamEndCapture(r->stream, &g);  // ends the capture on every joined stream
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return set_error(EXC_ERR_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
            e = hipGraphInstantiate(&r->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) return set_error(EXC_ERR_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
            for (size_t i = 0; i < r->cands.size(); ++i) r->graph_model_gen[i] = r->cands[i]->model_gen;
        }
        HIP_TRY(hipEventRecord(r->ev[0], r->stream));
        HIP_TRY(hipGraphLaunch(r->gexec, r->stream));
        HIP_TRY(hipEventRecord(r->ev[1], r->stream));
    } else {
        HIP_TRY(hipEventRecord(r->ev[0], r->stream));
        if (int rc = router_enqueue(r, r->ev[0])) return rc;
        HIP_TRY(hipEventRecord(r->ev[1], r->stream));
    }
    HIP_TRY(hipStreamSynchronize(r->stream));
    r->ms = 0;
    (void)hipEventElapsedTime(&r->ms, r->ev[0], r->ev[1]);
    for (size_t i = 0; i < r->kids.size(); ++i)
        if (r->active[i]) r->kids[i]->state = 3;
    r->state = 3;
    return EXC_OK;
}

// Best whole-run time over reps runs of the planned batch (after one warm run).
int router_time(exc_router* r, uint32_t reps, float* best)
{
    int rc = router_run(r);
    *best = 1e30f;
    for (uint32_t k = 0; k < reps && !rc; ++k)
        if (!(rc = router_run(r))) *best = std::min(*best, r->ms);
    return rc;
}

void router_set_persistent(exc_router* r, const std::vector<int>& p)
{
    for (size_t i = 0; i < r->kids.size(); ++i)
        if (r->kids[i]->persistent != p[i]) {
            r->kids[i]->persistent = p[i];
            drop_graph(r->kids[i]);
        }
    router_drop_graph(r);
}

}  // namespace

extern "C" int exc_router_create(exc_ctx* c, exc_index* const* cands, uint32_t n, exc_router** out)
{
    if (!c || !cands || !out || !n) return set_error(EXC_ERR_ARG, "exc_router_create: bad argument");
    if (n > EXC_ROUTER_MAX_CANDIDATES)
        return set_error(EXC_ERR_ARG, "exc_router_create: at most %d candidates", EXC_ROUTER_MAX_CANDIDATES);
    for (uint32_t i = 0; i < n; ++i) {
        const exc_index* ix = cands[i];
        if (!ix) return set_error(EXC_ERR_ARG, "exc_router_create: candidate %u is null", i);
        if (ix->codec == EXC_NEWPFD) return set_error(EXC_ERR_ARG, "NewPFD indexes are for decompression only");
        if (ix->ctx->device != c->device)
            return set_error(EXC_ERR_ARG, "candidate %u lives on device %d, the context on %d", i, ix->ctx->device,
                             c->device);
        if (ix->nterms != cands[0]->nterms || ix->n != cands[0]->n)
            return set_error(EXC_ERR_ARG, "candidate %u holds other lists than candidate 0", i);
    }
    HIP_TRY(hipSetDevice(c->device));
    auto* r = new exc_router();
    r->parent = c;
    r->cands.assign(cands, cands + n);
    r->active.assign(n, 0);
    r->join.assign(n, nullptr);
    r->graph_model_gen.assign(n, 0);
    int rc = EXC_OK;
    for (uint32_t i = 0; i < n && !rc; ++i) {
        exc_ctx* k = nullptr;
        if (!(rc = exc_ctx_create(c->device, c->budget, &k))) {
            k->search = c->search;
            k->tpe = c->tpe;
            k->coop_max = c->coop_max;
            k->persistent = c->persistent;
            k->timing = 0;
            r->kids.push_back(k);
        }
    }
    hipError_t e = rc ? hipSuccess : hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking);
    for (int j = 0; j < 2 && e == hipSuccess; ++j) e = hipEventCreate(&r->ev[j]);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&r->fork, hipEventDisableTiming);
    for (uint32_t i = 0; i < n && e == hipSuccess; ++i) e = hipEventCreateWithFlags(&r->join[i], hipEventDisableTiming);
    if (!rc && e != hipSuccess) rc = set_error(EXC_ERR_DEVICE, "exc_router_create: %s", hipGetErrorString(e));
    if (rc) {
        exc_router_destroy(r);
        return rc;
    }
    *out = r;
    return EXC_OK;
}

extern "C" void exc_router_destroy(exc_router* r)
{
    if (!r) return;
    (void)hipSetDevice(r->parent->device);
    if (r->stream) (void)hipStreamSynchronize(r->stream);
    router_drop_graph(r);
    for (exc_ctx* k : r->kids) exc_ctx_destroy(k);
    for (hipEvent_t e : r->ev)
        if (e) (void)hipEventDestroy(e);
    if (r->fork) (void)hipEventDestroy(r->fork);
    for (hipEvent_t e : r->join)
        if (e) (void)hipEventDestroy(e);
    if (r->stream) (void)hipStreamDestroy(r->stream);
    delete r;
}

extern "C" int exc_router_set_routes(exc_router* r, const uint32_t* route)
{
    if (!r || !route) return set_error(EXC_ERR_ARG, "exc_router_set_routes: null argument");
    for (int k = 0; k < kRatioClasses; ++k)
        if (route[k] >= r->cands.size())
            return set_error(EXC_ERR_ARG, "class %d: candidate %u of %zu", k, route[k], r->cands.size());
    std::memcpy(r->route, route, sizeof r->route);
    r->state = 0;  // the plan in hand was routed the old way
    router_drop_graph(r);
    return EXC_OK;
}

extern "C" int exc_router_set_option(exc_router* r, int key, int64_t v)
{
    if (!r) return set_error(EXC_ERR_ARG, "exc_router_set_option: null router");
    if (key == EXC_OPT_GRAPH) {
        r->graph = v != 0;
    } else if (key == EXC_OPT_TIMING) {
        // the router times every run itself; the candidates' own events stay off
    } else {
        for (exc_ctx* k : r->kids)
            if (int rc = exc_ctx_set_option(k, key, v)) return rc;
    }
    router_drop_graph(r);
    return EXC_OK;
}

extern "C" int exc_router_plan(exc_router* r, const uint32_t* qterms, const uint64_t* qoff, uint32_t nq)
{
    if (!r || !qoff || (nq && !qterms && qoff[nq])) return set_error(EXC_ERR_ARG, "exc_router_plan: null argument");
    return router_plan(r, qterms, qoff, nq);
}

extern "C" int exc_router_invoke(exc_router* r, float* ms)
{
    if (!r) return set_error(EXC_ERR_ARG, "exc_router_invoke: null router");
    int rc = router_run(r);
    if (!rc && ms) *ms = r->ms;
    return rc;
}

extern "C" int exc_router_fetch(exc_router* r, uint32_t* cum, uint32_t* res, uint64_t cap, uint64_t* nres)
{
    if (!r || (r->nq && !cum)) return set_error(EXC_ERR_ARG, "exc_router_fetch: null argument");
    if (r->state < 3) return set_error(EXC_ERR_STATE, "exc_router_fetch before exc_router_invoke");
    HIP_TRY(hipSetDevice(r->parent->device));
    const size_t n = r->kids.size();
    std::vector<std::vector<uint32_t>> kc(n), kr(n);
    uint64_t total = 0;
    for (size_t i = 0; i < n; ++i) {
        if (!r->active[i]) continue;
        exc_ctx* k = r->kids[i];
        kc[i].resize(k->nq);
        HIP_TRY(hipMemcpy(kc[i].data(), k->d_qcum.p, (size_t)k->nq * 4, hipMemcpyDeviceToHost));
        kr[i].resize(kc[i].back());
        if (!kr[i].empty()) HIP_TRY(hipMemcpy(kr[i].data(), k->d_result.p, kr[i].size() * 4, hipMemcpyDeviceToHost));
        total += kr[i].size();
    }
    if (nres) *nres = total;
    if (total > cap)
        return set_error(EXC_ERR_CAPACITY, "result buffer holds %llu, need %llu", (unsigned long long)cap,
                         (unsigned long long)total);
    if (total > 0xffffffffull) return set_error(EXC_ERR_CAPACITY, "more than 2^32 results in one batch");
    uint64_t o = 0;
    for (uint32_t q = 0; q < r->nq; ++q) {
        const uint32_t i = r->qcand[q], j = r->qidx[q];
        const uint32_t b = j ? kc[i][j - 1] : 0u, e = kc[i][j];
        if (e > b) std::memcpy(res + o, kr[i].data() + b, (size_t)(e - b) * 4);
        o += e - b;
        cum[q] = (uint32_t)o;
    }
    return EXC_OK;
}

extern "C" int exc_router_intersect(exc_router* r, const uint32_t* qterms, const uint64_t* qoff, uint32_t nq,
                                    uint32_t* cum, uint32_t* res, uint64_t cap, uint64_t* nres)
{
    if (int rc = exc_router_plan(r, qterms, qoff, nq)) return rc;
    if (int rc = router_run(r)) return rc;
    return exc_router_fetch(r, cum, res, cap, nres);
}

extern "C" int exc_router_autotune(exc_router* r, const uint32_t* qterms, const uint64_t* qoff, uint32_t nq,
                                   uint32_t reps, exc_route_table* out)
{
    if (!r || !qoff || (nq && !qterms) || !out) return set_error(EXC_ERR_ARG, "exc_router_autotune: bad argument");
    const uint32_t n = (uint32_t)r->cands.size();
    reps = std::max<uint32_t>(1, std::min<uint32_t>(reps, 50));
    exc_route_table T;
    std::memset(&T, 0, sizeof T);
    T.ncands = n;
    const bool had_graph = r->graph;
    r->graph = 0;
    router_drop_graph(r);
    // every candidate's knobs and launch mode on the sample
    std::vector<exc_tune_table> tt(n);
    std::vector<int> own(n);
    int rc = EXC_OK;
    for (uint32_t i = 0; i < n && !rc; ++i) {
        rc = exc_autotune(r->kids[i], r->cands[i], qterms, qoff, nq, reps, &tt[i]);
        own[i] = r->kids[i]->persistent;
        r->kids[i]->timing = 0;
        for (int k = 0; k < kRatioClasses; ++k) T.class_ms[i][k] = tt[i].ms[k];
    }
    uint32_t route[kRatioClasses] = {};
    if (!rc) {
        // each class to its fastest candidate; single-term queries probe
        // nothing and go where most of the sample goes
        std::vector<uint32_t> load(n, 0);
        for (int k = 0; k + 1 < kRatioClasses; ++k) {
            if (!tt[0].queries[k]) continue;
            for (uint32_t i = 1; i < n; ++i)
                if (T.class_ms[i][k] < T.class_ms[route[k]][k]) route[k] = i;
            load[route[k]] += tt[0].queries[k];
        }
        route[kRatioClasses - 1] = (uint32_t)(std::max_element(load.begin(), load.end()) - load.begin());
    }
    // the whole sample on every candidate alone, then routed, timed alike
    std::vector<uint32_t> solo(kRatioClasses);
    for (uint32_t i = 0; i < n && !rc; ++i) {
        std::fill(solo.begin(), solo.end(), i);
        std::memcpy(r->route, solo.data(), sizeof r->route);
        router_set_persistent(r, own);
        if (!(rc = router_plan(r, qterms, qoff, nq))) rc = router_time(r, reps, &T.single_ms[i]);
    }
    std::memcpy(r->route, route, sizeof r->route);
    T.routed_ms = 1e30f;
    for (int p = 0; p < 2 && !rc; ++p) {
        router_set_persistent(r, std::vector<int>(n, p));
        float ms = 0;
        if (!(rc = router_plan(r, qterms, qoff, nq)) && !(rc = router_time(r, reps, &ms)) && ms < T.routed_ms) {
            T.routed_ms = ms;
            T.persistent = (uint32_t)p;
        }
    }
    if (!rc) {
        const uint32_t best = (uint32_t)(std::min_element(T.single_ms, T.single_ms + n) - T.single_ms);
        if (T.single_ms[best] <= T.routed_ms) {
            for (auto& x : route) x = best;
            router_set_persistent(r, own);
            T.persistent = (uint32_t)own[best];
        } else {
            router_set_persistent(r, std::vector<int>(n, (int)T.persistent));
        }
        std::memcpy(r->route, route, sizeof r->route);
        std::memcpy(T.route, route, sizeof T.route);
    }
    r->graph = had_graph;
    r->state = 0;  // the last plan was the sample
    router_drop_graph(r);
    if (rc) return rc;
    *out = T;
    return EXC_OK;
}

// ---------------------------------------------------------------- serving

// Continuous-batching request loop.  The compressed index stays resident; the
// loop owns `nslots` batch contexts (one HIP stream each).  Arriving queries
// wait in a FIFO; whenever a slot is free the loop admits queries into a new
// batch until the batch holds max_batch queries or the device memory that the
// in-flight batches need (stage + result buffers, 8 B per shortest-list docID,
// plus per-query and per-unit records) would exceed the budget.  Batches on
// different slots overlap on the device; each finishing batch hands every
// query its own results and latency.
struct exc_server {
    // Pinned host buffers for batches' results, pooled (pinning is slow): a
    // batch's results are copied once, by DMA, into one buffer that its
    // requests share until the last of them is collected.
    struct Pinned {
        uint32_t* p = nullptr;
        size_t cap = 0;
    };
    struct Pool {
        std::mutex mu;
        std::vector<Pinned> free;
        ~Pool()
        {
            for (auto& b : free) (void)hipHostFree(b.p);
        }
    };
    std::shared_ptr<Pool> pool = std::make_shared<Pool>();
    struct Req {
        std::vector<uint32_t> terms;
        std::chrono::steady_clock::time_point arrival;
        std::shared_ptr<Pinned> buf;  // the batch's results
        uint64_t off = 0;             // this request's first result in buf
        uint64_t count = 0;
        double latency_ms = 0;
        uint64_t need = 0;
        bool done = false;
        int err = 0;
    };
    struct Slot {
        exc_ctx* ctx = nullptr;
        std::vector<uint64_t> reqs;  // tickets in the batch
        std::vector<uint32_t> qt;    // their terms, copied at admission
        std::vector<uint64_t> qo{0};
        uint64_t bytes = 0;
        bool busy = false;
        int phase = 0;               // 1: counts on their way, 2: results on their way
        hipEvent_t ev = nullptr;
        uint32_t* h_cum = nullptr;   // pinned
        size_t h_cap = 0;
        std::shared_ptr<Pinned> res;  // results of the batch in flight
    };
    exc_index* ix = nullptr;
    uint32_t max_batch = 1024;
    uint32_t flags = 0;
    uint64_t budget = 0, inflight = 0;
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv_req, cv_done;
    std::deque<uint64_t> queue;
    std::unordered_map<uint64_t, Req> reqs;  // open requests by ticket
    uint64_t next_ticket = 0;
    std::thread loop;
    std::atomic<bool> stop{false};
    // stats: written by the request loop, read by exc_server_stats from any thread
    std::atomic<uint64_t> batches{0}, queries_done{0}, max_inflight_batches{0}, overlap_launches{0};
};

namespace {

uint64_t query_need(exc_index* ix, const std::vector<uint32_t>& terms)
{
    uint32_t mn = 0;
    bool first = true;
    for (uint32_t t : terms) {
        if (t >= ix->nterms) return 0;
        if (first || ix->n[t] < mn) mn = ix->n[t];
        first = false;
    }
    return 8ull * mn + sizeof(QueryRec) + 4ull * terms.size() + 16ull * ((mn + kChunk - 1) / kChunk) + 64;
}

int slot_launch(exc_server* sv, exc_server::Slot& sl)
{
    exc_ctx* c = sl.ctx;
    int rc;
    if ((rc = plan_queries(c, sv->ix, sl.qt.data(), sl.qo.data(), (uint32_t)sl.reqs.size()))) return rc;
    if ((rc = upload_plan(c, false))) return rc;
    if ((rc = enqueue_invoke(c, false))) return rc;
    const size_t nq = sl.reqs.size();
    if (sl.h_cap < nq) {
        if (sl.h_cum) (void)hipHostFree(sl.h_cum);
        sl.h_cap = std::max<size_t>(nq, 2 * sl.h_cap);
        HIP_TRY(hipHostMalloc((void**)&sl.h_cum, sl.h_cap * 4, hipHostMallocDefault));
    }
    HIP_TRY(hipMemcpyAsync(sl.h_cum, c->d_qcum.p, nq * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipEventRecord(sl.ev, c->stream));
    sl.phase = 1;
    c->state = 3;
    return EXC_OK;
}

// Counts are on the host: start the results' copy into the slot's pinned
// buffer on its own stream (phase 2).  Returns false when there is nothing to
// copy (counts-only server or empty results).
int slot_copy_results(exc_server* sv, exc_server::Slot& sl, bool& started)
{
    started = false;
    const size_t nq = sl.reqs.size();
    const uint64_t total = nq ? sl.h_cum[nq - 1] : 0;
    if ((sv->flags & EXC_SERVER_COUNTS) || total == 0) return EXC_OK;
    exc_server::Pinned b;
    {
        std::lock_guard<std::mutex> g(sv->pool->mu);
        auto& fl = sv->pool->free;
        for (size_t i = 0; i < fl.size(); ++i)
            if (fl[i].cap >= total) {
                b = fl[i];
                fl.erase(fl.begin() + (long)i);
                break;
            }
    }
    if (!b.p) {
        // twice the batch (at least 16 Mi results): a pinned allocation stalls the loop
        b.cap = std::max<size_t>(2 * total, 1u << 24);
        HIP_TRY(hipHostMalloc((void**)&b.p, b.cap * 4, hipHostMallocDefault));
    }
    std::weak_ptr<exc_server::Pool> wp = sv->pool;
    sl.res = std::shared_ptr<exc_server::Pinned>(new exc_server::Pinned(b), [wp](exc_server::Pinned* x) {
        if (auto pl = wp.lock()) {
            std::lock_guard<std::mutex> g(pl->mu);
            pl->free.push_back(*x);
        } else {
            (void)hipHostFree(x->p);
        }
        delete x;
    });
    HIP_TRY(hipMemcpyAsync(sl.res->p, sl.ctx->d_result.p, total * 4, hipMemcpyDeviceToHost, sl.ctx->stream));
    HIP_TRY(hipEventRecord(sl.ev, sl.ctx->stream));
    sl.phase = 2;
    started = true;
    return EXC_OK;
}

void slot_finish(exc_server* sv, exc_server::Slot& sl, int err)
{
    const size_t nq = sl.reqs.size();
    const bool with_res = !err && !(sv->flags & EXC_SERVER_COUNTS) && sl.phase == 2;
    const auto now = std::chrono::steady_clock::now();
    std::lock_guard<std::mutex> g(sv->mu);
    for (size_t i = 0; i < nq; ++i) {
        auto& r = sv->reqs.at(sl.reqs[i]);
        r.err = err;
        if (!err) {
            const uint32_t b = i ? sl.h_cum[i - 1] : 0, e = sl.h_cum[i];
            r.count = e - b;
            if (with_res) {
                r.buf = sl.res;
                r.off = b;
            }
        }
        r.latency_ms = std::chrono::duration<double, std::milli>(now - r.arrival).count();
        r.done = true;
    }
    sv->queries_done += nq;
    sv->inflight -= sl.bytes;
    sl.reqs.clear();
    sl.qt.clear();
    sl.qo.assign(1, 0);
    sl.bytes = 0;
    sl.busy = false;
    sl.phase = 0;
    sl.res.reset();
    sv->cv_done.notify_all();
}

void server_loop(exc_server* sv)
{
    (void)hipSetDevice(sv->slots[0].ctx->device);
    for (;;) {
        bool progressed = false;
        // 1. move batches along: counts in -> results copy, results in -> retire
        for (auto& sl : sv->slots) {
            if (sl.busy && hipEventQuery(sl.ev) == hipSuccess) {
                bool started = false;
                int rc = sl.phase == 1 ? slot_copy_results(sv, sl, started) : EXC_OK;
                if (rc || !started) slot_finish(sv, sl, rc);
                progressed = true;
            }
        }
        // 2. admit waiting queries into free slots
        for (auto& sl : sv->slots) {
            if (sl.busy) continue;
            {
                std::unique_lock<std::mutex> lk(sv->mu);
                while (!sv->queue.empty() && sl.reqs.size() < sv->max_batch) {
                    const uint64_t tk = sv->queue.front();
                    const auto& r = sv->reqs.at(tk);
                    if (sv->budget && sv->inflight + sl.bytes + r.need > sv->budget && (sl.bytes || sv->inflight)) break;
                    sv->queue.pop_front();
                    sl.reqs.push_back(tk);
                    sl.qt.insert(sl.qt.end(), r.terms.begin(), r.terms.end());
                    sl.qo.push_back(sl.qt.size());
                    sl.bytes += r.need;
                }
                if (sl.reqs.empty()) break;
                sv->inflight += sl.bytes;
            }
            uint32_t busy_now = 0;
            for (auto& o : sv->slots) busy_now += o.busy;
            if (busy_now) ++sv->overlap_launches;
            if (busy_now + 1 > sv->max_inflight_batches.load()) sv->max_inflight_batches.store(busy_now + 1);
            const int rc = slot_launch(sv, sl);
            ++sv->batches;
            sl.busy = true;
            if (rc) slot_finish(sv, sl, rc);
            progressed = true;
        }
        if (sv->stop.load()) {
            bool any = false;
            for (auto& sl : sv->slots) any = any || sl.busy;
            std::lock_guard<std::mutex> g(sv->mu);
            if (!any && sv->queue.empty()) break;
        }
        if (!progressed) {
            bool any = false;
            for (auto& sl : sv->slots) any = any || sl.busy;
            std::unique_lock<std::mutex> lk(sv->mu);
            if (!any && sv->queue.empty() && !sv->stop.load())
                sv->cv_req.wait_for(lk, std::chrono::milliseconds(5));
            else if (any)
                std::this_thread::yield();
        }
    }
}

}  // namespace

extern "C" int exc_server_create(exc_ctx* c, exc_index* ix, uint32_t nslots, uint32_t max_batch, uint64_t budget,
                                 uint32_t flags, exc_server** out)
{
    if (!c || !ix || !out || nslots == 0 || nslots > 16 || max_batch == 0)
        return set_error(EXC_ERR_ARG, "exc_server_create: bad argument");
    if (ix->codec == EXC_NEWPFD) return set_error(EXC_ERR_ARG, "NewPFD indexes are for decompression only");
    auto* sv = new exc_server();
    sv->ix = ix;
    sv->max_batch = max_batch;
    sv->budget = budget;
    sv->flags = flags;
    sv->slots.resize(nslots);
    for (auto& sl : sv->slots) {
        int rc = exc_ctx_create(c->device, 0, &sl.ctx);
        if (rc) {
            for (auto& o : sv->slots)
                if (o.ctx) exc_ctx_destroy(o.ctx);
            delete sv;
            return rc;
        }
        sl.ctx->search = c->search;
        sl.ctx->tpe = c->tpe;
        sl.ctx->persistent = c->persistent;
        sl.ctx->coop_max = c->coop_max;
        sl.ctx->tuned = c->tuned;
        sl.ctx->tune = c->tune;
        sl.ctx->timing = 0;
        (void)hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming);
    }
    sv->loop = std::thread(server_loop, sv);
    *out = sv;
    return EXC_OK;
}

extern "C" int exc_server_submit(exc_server* sv, const uint32_t* terms, uint32_t nterms, uint64_t* ticket)
{
    if (!sv || (nterms && !terms) || !ticket) return set_error(EXC_ERR_ARG, "exc_server_submit: bad argument");
    exc_server::Req r;
    r.terms.assign(terms, terms + nterms);
    for (uint32_t t : r.terms)
        if (t >= sv->ix->nterms) return set_error(EXC_ERR_ARG, "query term %u out of range", t);
    r.need = query_need(sv->ix, r.terms);
    if (sv->budget && r.need > sv->budget)
        return set_error(EXC_ERR_BUDGET, "query needs %llu bytes, more than the whole budget",
                         (unsigned long long)r.need);
    r.arrival = std::chrono::steady_clock::now();
    {
        std::lock_guard<std::mutex> g(sv->mu);
        *ticket = sv->next_ticket++;
        sv->reqs.emplace(*ticket, std::move(r));
        sv->queue.push_back(*ticket);
    }
    sv->cv_req.notify_one();
    return EXC_OK;
}

extern "C" int exc_server_wait(exc_server* sv, uint64_t ticket, uint32_t* results, uint64_t cap, uint64_t* n,
                               double* latency_ms)
{
    if (!sv) return set_error(EXC_ERR_ARG, "exc_server_wait: null server");
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(ticket);
    if (it == sv->reqs.end()) return set_error(EXC_ERR_ARG, "unknown or already collected ticket");
    // the wait releases the mutex and a concurrent submit may rehash the map:
    // iterators die with a rehash, references to the element do not
    exc_server::Req& r = it->second;
    sv->cv_done.wait(lk, [&] { return r.done; });
    if (n) *n = r.count;
    if (latency_ms) *latency_ms = r.latency_ms;
    if (r.err) {
        const int e = r.err;
        sv->reqs.erase(ticket);
        return set_error(e, "query failed on the device");
    }
    // a too-small buffer keeps the request so the caller can retry with *n
    if (!(sv->flags & EXC_SERVER_COUNTS) && r.count > cap)
        return set_error(EXC_ERR_CAPACITY, "result buffer too small");
    // the copy runs outside the lock (the server loop and the submitters take
    // it), on a reference that keeps the batch's pinned buffer alive
    const std::shared_ptr<exc_server::Pinned> buf = r.buf;
    const uint64_t off = r.off, count = r.count;
    sv->reqs.erase(ticket);
    lk.unlock();
    if (results && buf && count) std::memcpy(results, buf->p + off, count * 4);
    return EXC_OK;
}

extern "C" int exc_server_replay(exc_server* sv, const uint32_t* qterms, const uint64_t* qoff, const double* arrival_s,
                                 uint32_t nq, double* lat_ms, uint32_t* counts, uint32_t* results, uint64_t res_cap)
{
    if (!sv || !qoff || !arrival_s || (qoff[nq] && !qterms))
        return set_error(EXC_ERR_ARG, "exc_server_replay: bad argument");
    for (uint32_t q = 0; q < nq; ++q)
        if (qoff[q + 1] < qoff[q]) return set_error(EXC_ERR_ARG, "exc_server_replay: query %u: bad offsets", q);
    for (uint64_t k = 0; k < qoff[nq]; ++k)
        if (qterms[k] >= sv->ix->nterms) return set_error(EXC_ERR_ARG, "query term %u out of range", qterms[k]);
    std::vector<uint64_t> tickets(nq);
    std::atomic<uint32_t> submitted{0};
    // Completed queries are collected (results copied out, buffers released)
    // while later ones are still being submitted.
    int crc = EXC_OK, cq = 0;
    std::thread collector([&] {
        uint64_t pos = 0;
        for (uint32_t q = 0; q < nq; ++q) {
            while (submitted.load(std::memory_order_acquire) <= q) std::this_thread::yield();
            std::unique_lock<std::mutex> lk(sv->mu);
            // a reference, not the iterator, across the wait (submits rehash)
            exc_server::Req& r = sv->reqs.at(tickets[q]);
            sv->cv_done.wait(lk, [&] { return r.done; });
            if (r.err && !crc) {
                crc = r.err;
                cq = (int)q;
            }
            if (lat_ms) lat_ms[q] = r.latency_ms;
            if (counts) counts[q] = (uint32_t)r.count;
            // the results are copied after the lock is released (the server
            // loop and the submitter take it); the reference keeps the batch's
            // pinned buffer alive
            const std::shared_ptr<exc_server::Pinned> buf = r.buf;
            const uint64_t off = r.off, count = r.count;
            sv->reqs.erase(tickets[q]);
            lk.unlock();
            if (results && buf && count) {
                if (pos + count > res_cap) {
                    if (!crc) {
                        crc = EXC_ERR_CAPACITY;
                        cq = -1;
                    }
                } else {
                    std::memcpy(results + pos, buf->p + off, count * 4);
                }
            }
            pos += buf ? count : 0;
        }
    });
    const auto t0 = std::chrono::steady_clock::now();
    for (uint32_t q = 0; q < nq; ++q) {
        const auto at = t0 + std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                                 std::chrono::duration<double>(arrival_s[q]));
        while (std::chrono::steady_clock::now() < at) std::this_thread::yield();
        exc_server::Req r;
        r.terms.assign(qterms + qoff[q], qterms + qoff[q + 1]);
        r.need = query_need(sv->ix, r.terms);
        r.arrival = at;  // latency counts from the scheduled arrival
        {
            std::lock_guard<std::mutex> g(sv->mu);
            tickets[q] = sv->next_ticket++;
            sv->reqs.emplace(tickets[q], std::move(r));
            sv->queue.push_back(tickets[q]);
        }
        submitted.store(q + 1, std::memory_order_release);
        sv->cv_req.notify_one();
    }
    collector.join();
    if (crc == EXC_ERR_CAPACITY && cq < 0) return set_error(crc, "replay result buffer too small");
    if (crc) return set_error(crc, "query %d failed on the device", cq);
    return EXC_OK;
}

extern "C" int exc_server_stats(exc_server* sv, exc_server_info* st)
{
    if (!sv || !st) return set_error(EXC_ERR_ARG, "exc_server_stats: bad argument");
    std::lock_guard<std::mutex> g(sv->mu);
    st->batches = sv->batches.load();
    st->queries = sv->queries_done.load();
    st->max_inflight_batches = sv->max_inflight_batches.load();
    st->overlapped_launches = sv->overlap_launches.load();
    return EXC_OK;
}

extern "C" void exc_server_destroy(exc_server* sv)
{
    if (!sv) return;
    sv->stop.store(true);
    sv->cv_req.notify_all();
    if (sv->loop.joinable()) sv->loop.join();
    for (auto& sl : sv->slots) {
        if (sl.ev) (void)hipEventDestroy(sl.ev);
        if (sl.h_cum) (void)hipHostFree(sl.h_cum);
        exc_ctx_destroy(sl.ctx);
    }
    delete sv;
}
