// gpu_conjunction_scorer.h -- CLucene host glue for the GPU conjunction path.
//
// CLucene answers a conjunctive query with ConjunctionScorer
// (src/core/CLucene/search/ConjunctionScorer.cpp:52-104): it leapfrogs its
// sub-scorers with next()/skipTo() until they agree on a docID, and scores a
// match as the sum of the sub-scorers' scores times coord(n, n) (:123-129).
// BooleanScorer2 builds one for required clauses
// (src/core/CLucene/search/BooleanScorer2.cpp:455-465).
//
// Here the intersection runs on the GPU instead, for many queries at once:
//
//   ConjunctionBatch      plans, runs and collects a batch of term-id queries
//                         with the library's stage functions (exc_plan,
//                         exc_kernelInvoke, exc_dtohTransfer); results of
//                         query q are ascending docIDs.
//   ConjunctionAdapter<B> the Scorer contract over one query's results:
//                         next(), doc(), skipTo(target) with the semantics of
//                         ConjunctionScorer, and score() = coord * the sum of
//                         the sub-scorers' scores at doc(), each sub-scorer
//                         moved there with skipTo() only when a score is
//                         asked for.  B is the Scorer base class: for CLucene
//                         lucene::search::Scorer, whose remaining virtuals
//                         (explain, toString) the integrating class adds.
//
// Header-only C++11 over the C ABI in exclucene_amd.h; errors throw
// exc_clucene::Error carrying the library status and message.
#ifndef EXCLUCENE_AMD_GPU_CONJUNCTION_SCORER_H
#define EXCLUCENE_AMD_GPU_CONJUNCTION_SCORER_H

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "exclucene_amd.h"

namespace exc_clucene {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& what) : std::runtime_error(what), code(c) {}
};

inline void check(int rc, const char* where)
{
    if (rc != EXC_OK) throw Error(rc, std::string(where) + ": " + exc_last_error());
}

// The results of a batch of conjunctive queries, intersected on the device
// in one plan / invoke / transfer sequence (one ctx: one host thread).
class ConjunctionBatch {
public:
    ConjunctionBatch(exc_ctx* ctx, exc_index* ix, const std::vector<std::vector<uint32_t> >& queries)
    {
        std::vector<uint32_t> terms;
        std::vector<uint64_t> qoff(1, 0);
        for (size_t q = 0; q < queries.size(); ++q) {
            terms.insert(terms.end(), queries[q].begin(), queries[q].end());
            qoff.push_back(terms.size());
        }
        const uint32_t nq = (uint32_t)queries.size();
        exc_batchInfo bi;
        check(exc_plan(ctx, ix, terms.data(), qoff.data(), nq, &bi), "exc_plan");
        check(exc_kernelInvoke(ctx, &bi), "exc_kernelInvoke");
        std::vector<uint32_t> cum(nq);
        // a query's results never outnumber its shortest list
        docs_.resize(std::max<uint64_t>(1, bi.ucntShortest_sum));
        check(exc_dtohTransfer(ctx, &bi, cum.data(), docs_.data(), docs_.size()), "exc_dtohTransfer");
        start_.assign(nq + 1, 0);
        for (uint32_t q = 0; q < nq; ++q) start_[q + 1] = cum[q];
        docs_.resize(nq ? cum[nq - 1] : 0);
    }
    size_t size() const { return start_.empty() ? 0 : start_.size() - 1; }
    const uint32_t* begin(size_t q) const { return docs_.data() + start_[q]; }
    const uint32_t* end(size_t q) const { return docs_.data() + start_[q + 1]; }
    uint64_t count(size_t q) const { return start_[q + 1] - start_[q]; }

private:
    std::vector<uint32_t> docs_;
    std::vector<uint64_t> start_;
};

// Scorer over one query's GPU results.  Sub is the sub-scorer type (the
// Scorer base itself for CLucene); coord is Similarity::coord(n, n).
template <class Base, class Sub = Base>
class ConjunctionAdapter : public Base {
public:
    template <class Sim>
    ConjunctionAdapter(Sim* similarity, const ConjunctionBatch& batch, size_t q, std::vector<Sub*> subs)
        : Base(similarity), b_(batch.begin(q)), e_(batch.end(q)), p_(0), subs_(std::move(subs)),
          coord_((float)similarity->coord((int32_t)subs_.size(), (int32_t)subs_.size()))
    {
    }

    // the lowest result above doc(), or the first one on the first call
    bool next()
    {
        p_ = p_ ? std::min(p_ + 1, e_) : b_;
        return p_ < e_;
    }

    // the current result; -1 before the first next()/skipTo() (the
    // reference's lastDoc), unspecified once exhausted
    int32_t doc() const { return (p_ && p_ < e_) ? (int32_t)*p_ : -1; }

    // the lowest result >= target; like the reference, undefined for a target
    // <= doc() (here: the position does not move back)
    bool skipTo(int32_t target)
    {
        const uint32_t* from = p_ ? p_ : b_;
        p_ = target <= 0 ? from : std::lower_bound(from, e_, (uint32_t)target);
        return p_ < e_;
    }

    using Base::score;  // CLucene's score(HitCollector*) overloads stay visible

    // coord * sum of the sub-scorers' scores at doc()
    float score()
    {
        const int32_t d = doc();
        float sum = 0.0f;
        for (size_t i = 0; i < subs_.size(); ++i) {
            if (subs_[i]->doc() < d) subs_[i]->skipTo(d);
            sum += subs_[i]->score();
        }
        return sum * coord_;
    }

private:
    const uint32_t *b_, *e_, *p_;
    std::vector<Sub*> subs_;
    float coord_;
};

}  // namespace exc_clucene

#endif
