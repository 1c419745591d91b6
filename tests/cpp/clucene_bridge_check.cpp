// clucene_bridge_check.cpp -- GPU check of exc_clucene::ConjunctionAdapter
// against a CPU restatement of CLucene's ConjunctionScorer iteration
// (src/core/CLucene/search/ConjunctionScorer.cpp:52-121) over term scorers
// with TermScorer's skipTo behaviour (advance at least one position, then to
// the first docID >= target).  Seeded lists are built into a SegLRC-128
// index, a batch of 1-4 term queries runs through ConjunctionBatch, and for
// every query both scorers are driven by the same sequence of next() and
// skipTo(doc() + k) calls; returned flags, doc() and score() must agree
// exactly.  Mock term scores are multiples of 1/8 below 1, so sums of up to
// four are exact in any order and the comparison is bit-exact.
// Exit status 0 when everything agrees.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "exclucene_amd.h"
#include "gpu_conjunction_scorer.h"

namespace {

struct Similarity {
    float coord(int32_t overlap, int32_t max_overlap) const { return (float)overlap / (float)max_overlap; }
};

struct Scorer {
    explicit Scorer(const Similarity*) {}
    virtual ~Scorer() {}
    virtual bool next() = 0;
    virtual int32_t doc() const = 0;
    virtual bool skipTo(int32_t target) = 0;
    virtual float score() = 0;
};

// one term's postings
struct TermScorer : Scorer {
    const std::vector<uint32_t>* docs;
    uint32_t term;
    int64_t pos = -1;
    TermScorer(const Similarity* s, const std::vector<uint32_t>* d, uint32_t t) : Scorer(s), docs(d), term(t) {}
    bool next() override { return ++pos < (int64_t)docs->size(); }
    int32_t doc() const override
    {
        return pos < 0 ? -1 : pos < (int64_t)docs->size() ? (int32_t)(*docs)[pos] : INT32_MAX;
    }
    bool skipTo(int32_t target) override
    {
        for (++pos; pos < (int64_t)docs->size(); ++pos)
            if ((int32_t)(*docs)[pos] >= target) return true;
        return false;
    }
    float score() override { return (float)(((uint32_t)doc() * 7u + term * 3u) & 7u) / 8.0f; }
};

// ConjunctionScorer's iteration: init() advances every scorer, orders them by
// doc and leapfrogs (doNext); later calls move the last scorer and leapfrog.
struct Leapfrog : Scorer {
    std::vector<Scorer*> s;
    bool first = true, more = false;
    int32_t last = -1;
    float coord;
    Leapfrog(const Similarity* sim, std::vector<Scorer*> v)
        : Scorer(sim), s(std::move(v)), coord(sim->coord((int32_t)s.size(), (int32_t)s.size()))
    {
    }
    bool leap()
    {
        size_t f = 0;
        Scorer* tail = s.back();
        while (more && s[f]->doc() < (last = tail->doc())) {
            more = s[f]->skipTo(last);
            tail = s[f];
            f = (f == s.size() - 1) ? 0 : f + 1;
        }
        return more;
    }
    bool init(int32_t target)
    {
        first = false;
        for (Scorer* x : s) {
            more = target == 0 ? x->next() : x->skipTo(target);
            if (!more) return false;
        }
        std::sort(s.begin(), s.end(), [](const Scorer* a, const Scorer* b) { return a->doc() < b->doc(); });
        leap();
        const int64_t end = (int64_t)s.size() - 2;
        for (int64_t i = 0; i < (end >> 1); ++i) std::swap(s[i], s[end - i]);
        return more;
    }
    bool next() override
    {
        if (first)
            init(0);
        else if (more)
            more = s.back()->next();
        return leap();
    }
    bool skipTo(int32_t target) override
    {
        if (first) return init(target);
        if (more) more = s.back()->skipTo(target);
        return leap();
    }
    int32_t doc() const override { return last; }
    float score() override
    {
        float sum = 0.0f;
        for (Scorer* x : s) sum += x->score();
        return sum * coord;
    }
};

std::vector<uint32_t> sorted_sample(std::mt19937_64& g, uint32_t n, uint32_t universe)
{
    std::vector<uint32_t> v;
    std::uniform_int_distribution<uint32_t> d(0, universe - 1);
    while (v.size() < n) {
        for (uint32_t i = (uint32_t)v.size(); i < n; ++i) v.push_back(d(g));
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    }
    return v;
}

int fail(const char* what)
{
    std::fprintf(stderr, "clucene bridge: %s: %s\n", what, exc_last_error());
    return 2;
}

}  // namespace

int main()
{
    const uint32_t universe = 1u << 22, nterms = 48, nq = 400;
    std::mt19937_64 g(8);
    std::vector<std::vector<uint32_t> > lists(nterms);
    std::vector<uint32_t> docs;
    std::vector<uint64_t> off(1, 0);
    std::uniform_real_distribution<double> ld(0.0, 5.5);
    for (uint32_t t = 0; t < nterms; ++t) {
        const uint32_t n = t == 0 ? 1u : t == 1 ? universe / 2 : (uint32_t)std::pow(10.0, ld(g));
        lists[t] = sorted_sample(g, n, universe);
        docs.insert(docs.end(), lists[t].begin(), lists[t].end());
        off.push_back(docs.size());
    }
    std::printf("clucene bridge: %u lists, %zu postings\n", nterms, docs.size());
    std::fflush(stdout);

    exc_ctx* ctx = nullptr;
    exc_built* built = nullptr;
    exc_index* ix = nullptr;
    exc_index_host view;
    if (exc_ctx_create(0, 0, &ctx)) return fail("exc_ctx_create");
    if (exc_build_index(EXC_SEGLRC, 128, 0.0f, 256, 0, docs.data(), off.data(), nterms, 4, &built))
        return fail("exc_build_index");
    if (exc_built_view(built, &view) || exc_index_upload(ctx, &view, &ix)) return fail("upload");

    std::vector<std::vector<uint32_t> > queries(nq);
    std::uniform_int_distribution<uint32_t> td(0, nterms - 1), kd(1, 4);
    for (auto& q : queries) {
        const uint32_t k = kd(g);
        while (q.size() < k) {
            const uint32_t t = td(g);
            if (std::find(q.begin(), q.end(), t) == q.end()) q.push_back(t);
        }
    }
    const Similarity sim;
    uint64_t results = 0, calls = 0, bad = 0;
    try {
        exc_clucene::ConjunctionBatch batch(ctx, ix, queries);
        for (uint32_t q = 0; q < nq; ++q) {
            results += batch.count(q);
            std::vector<TermScorer> a, b;
            a.reserve(queries[q].size());
            b.reserve(queries[q].size());
            for (uint32_t t : queries[q]) {
                a.emplace_back(&sim, &lists[t], t);
                b.emplace_back(&sim, &lists[t], t);
            }
            std::vector<Scorer*> pa, pb;
            for (auto& x : a) pa.push_back(&x);
            for (auto& x : b) pb.push_back(&x);
            Leapfrog ref(&sim, pa);
            exc_clucene::ConjunctionAdapter<Scorer> gpu(&sim, batch, q, pb);
            // even queries walk with next() only; odd ones mix in skips
            std::uniform_int_distribution<int32_t> skip(1, 1 << 16);
            for (bool more = true; more;) {
                const bool use_skip = (q & 1) && (calls % 3 == 2) && ref.doc() >= 0;
                const int32_t target = use_skip ? ref.doc() + skip(g) : 0;
                const bool r1 = use_skip ? ref.skipTo(target) : ref.next();
                const bool r2 = use_skip ? gpu.skipTo(target) : gpu.next();
                ++calls;
                if (r1 != r2 || (r1 && (ref.doc() != gpu.doc() || ref.score() != gpu.score()))) {
                    if (bad++ < 5)
                        std::fprintf(stderr, "query %u call %llu: ref (%d, %d, %g) gpu (%d, %d, %g)\n", q,
                                     (unsigned long long)calls, (int)r1, r1 ? ref.doc() : -1,
                                     r1 ? ref.score() : 0.0, (int)r2, r2 ? gpu.doc() : -1, r2 ? gpu.score() : 0.0);
                    break;
                }
                more = r1;
            }
        }
    } catch (const exc_clucene::Error& e) {
        std::fprintf(stderr, "clucene bridge: %s\n", e.what());
        return 2;
    }
    exc_index_free(ix);
    exc_built_free(built);
    exc_ctx_destroy(ctx);
    std::printf("clucene bridge: %u queries, %llu results, %llu scorer calls, %s\n", nq,
                (unsigned long long)results, (unsigned long long)calls, bad ? "MISMATCH" : "all equal");
    return bad ? 1 : 0;
}
