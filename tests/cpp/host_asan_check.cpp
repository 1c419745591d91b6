// host_asan_check.cpp -- the host index builder and file writer under
// AddressSanitizer and UndefinedBehaviorSanitizer (no GPU needed).
//
// Built by tests/cpp/Makefile with g++ -fsanitize=address,undefined from the
// library's host-only sources (builder.cpp, errors.cpp) and the oracle
// (oracle.c), so every builder path runs instrumented: each codec over lists
// with the edge lengths (empty, 1, a segment +- 1, long), fixed and adaptive
// hash widths, the files of every codec written out, the refusals of bad
// input, and the synthetic-list generator.  Every built index is decoded and
// queried back through the oracle, against std::set_intersection of the
// source lists.  Exits non-zero on a mismatch; a sanitizer report aborts.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include <sys/stat.h>
#include <unistd.h>

#include "exclucene_amd.h"
extern "C" {
#include "oracle.h"
}

namespace {

int failures = 0;

void check(bool ok, const char* what, int codec)
{
    if (!ok) {
        std::fprintf(stderr, "FAIL codec %d: %s\n", codec, what);
        ++failures;
    }
}

std::vector<std::vector<uint32_t>> make_lists(uint32_t seed)
{
    std::mt19937_64 g(seed);
    std::vector<std::vector<uint32_t>> out;
    const uint32_t lens[] = {0, 1, 2, 127, 128, 129, 255, 256, 257, 1000, 5000, 40000};
    for (uint32_t n : lens) {
        std::vector<uint32_t> L;
        std::uniform_int_distribution<uint32_t> d(0, (1u << 25) - 1);
        while (L.size() < n) {
            L.push_back(d(g));
            if (L.size() == n) {
                std::sort(L.begin(), L.end());
                L.erase(std::unique(L.begin(), L.end()), L.end());
            }
        }
        out.push_back(L);
    }
    std::vector<uint32_t> dense(3000);  // a dense run and extremes
    for (uint32_t i = 0; i < dense.size(); ++i) dense[i] = 1000 + i;
    out.push_back(dense);
    out.push_back({0, 5, 6, 7, 0x7fffffffu});
    return out;
}

orc_index view_to_oracle(const exc_index_host& h, std::vector<uint32_t>& ucnt, std::vector<uint64_t>& uw)
{
    orc_index o;
    std::memset(&o, 0, sizeof o);
    o.codec = h.codec;
    o.seg = h.seg_size;
    o.nterms = h.nterms;
    o.total_words = h.ind2_words;
    ucnt.resize(h.nterms);
    uw.resize(h.nterms);
    for (uint32_t t = 0; t < h.nterms; ++t) {
        ucnt[t] = h.ind1[t].m_urlcount;
        uw[t] = h.ind1[t].m_off / 4;
    }
    o.ucnt = ucnt.data();
    o.uwlist = uw.data();
    o.segnum = h.seg_num;
    o.words = h.ind2;
    o.seg_off = h.seg_offset;
    o.seg_head = h.seg_head;
    o.seg_median = h.seg_median;
    o.lrc = (const orc_lrc_info*)h.lrc_info;
    o.fit = (const orc_seg_info*)h.seg_info;
    o.base = h.seg_base;
    o.hash_off = h.hash_offset;
    o.hash_info = h.hash_info;
    return o;
}

void run_codec(const std::vector<std::vector<uint32_t>>& lists, int codec, uint32_t seg, float frac, uint32_t shift,
               const std::string& dir)
{
    std::vector<uint32_t> docs;
    std::vector<uint64_t> off{0};
    for (const auto& L : lists) {
        docs.insert(docs.end(), L.begin(), L.end());
        off.push_back(docs.size());
    }
    exc_built* b = nullptr;
    if (exc_build_index(codec, seg, frac, 256, shift, docs.data(), off.data(), (uint32_t)lists.size(), 3, &b)) {
        std::fprintf(stderr, "build codec %d: %s\n", codec, exc_last_error());
        ++failures;
        return;
    }
    exc_index_host h;
    check(exc_built_view(b, &h) == EXC_OK, "view", codec);
    check(exc_index_save_files(&h, dir.c_str(), "idx") == EXC_OK, "save", codec);
    if (codec == EXC_RAW || codec == EXC_RAW_HS) {
        // the docIDs themselves
        for (uint32_t t = 0; t < h.nterms; ++t)
            check(std::equal(lists[t].begin(), lists[t].end(), h.ind2 + h.ind1[t].m_off / 4), "raw words", codec);
        exc_built_free(b);
        return;
    }
    std::vector<uint32_t> ucnt;
    std::vector<uint64_t> uw;
    orc_index o = view_to_oracle(h, ucnt, uw);
    std::vector<uint32_t> dec;
    for (uint32_t t = 0; t < h.nterms; ++t) {
        dec.assign(lists[t].size() + 1, 0);
        check(orc_decode_term(&o, t, dec.data()) == 0, "decode", codec);
        check(std::equal(lists[t].begin(), lists[t].end(), dec.begin()), "round trip", codec);
    }
    if (codec == EXC_NEWPFD) {  // decompression only, as in the reference
        exc_built_free(b);
        return;
    }
    // every pair of lists as a query, against std::set_intersection
    std::vector<uint32_t> qterms;
    std::vector<uint64_t> qoff{0};
    std::vector<uint32_t> want;
    std::vector<uint32_t> want_cum;
    for (uint32_t a = 0; a < h.nterms; ++a)
        for (uint32_t c = 0; c < h.nterms; ++c) {
            qterms.push_back(a);
            qterms.push_back(c);
            qoff.push_back(qterms.size());
            std::set_intersection(lists[a].begin(), lists[a].end(), lists[c].begin(), lists[c].end(),
                                  std::back_inserter(want));
            want_cum.push_back((uint32_t)want.size());
        }
    const uint32_t nq = (uint32_t)(qoff.size() - 1);
    std::vector<uint32_t> cum(nq), res(want.size() + 1);
    check(orc_eval_queries(&o, qterms.data(), qoff.data(), nq, 2, cum.data(), res.data(), res.size(), nullptr) == 0,
          "eval", codec);
    check(cum == want_cum, "counts", codec);
    check(std::equal(want.begin(), want.end(), res.begin()), "results", codec);
    exc_built_free(b);
}

}  // namespace

int main()
{
    char tmpl[] = "/tmp/exc_asan_XXXXXX";
    const char* dir = mkdtemp(tmpl);
    if (!dir) return 2;
    const auto lists = make_lists(7);
    run_codec(lists, EXC_SEGLRC, 128, 0.0f, 0, dir);
    run_codec(lists, EXC_SEGLRC, 256, 0.0f, 0, dir);
    run_codec(lists, EXC_LRC, 256, 0.0f, 0, dir);
    run_codec(lists, EXC_PARAPFD, 128, 0.1f, 0, dir);
    run_codec(lists, EXC_PARAPFD, 256, 0.5f, 0, dir);
    run_codec(lists, EXC_NEWPFD, 128, 0.1f, 0, dir);
    run_codec(lists, EXC_HS_SEGLRC, 128, 0.0f, 0, dir);
    run_codec(lists, EXC_HS_SEGLRC, 128, 0.0f, 13, dir);
    run_codec(lists, EXC_RAW, 128, 0.0f, 0, dir);
    run_codec(lists, EXC_RAW_HS, 128, 0.0f, 0, dir);
    run_codec(lists, EXC_RAW_HS, 128, 0.0f, 13, dir);
    // refusals: unsorted list, docID >= 2^31, bad segment size, bad shift
    {
        const uint32_t bad[] = {5, 3, 9};
        const uint64_t off[] = {0, 3};
        exc_built* b = nullptr;
        check(exc_build_index(EXC_SEGLRC, 128, 0.0f, 256, 0, bad, off, 1, 1, &b) == EXC_ERR_ARG, "unsorted", 1);
        const uint32_t big[] = {1, 0x80000000u};
        const uint64_t off2[] = {0, 2};
        check(exc_build_index(EXC_SEGLRC, 128, 0.0f, 256, 0, big, off2, 1, 1, &b) == EXC_ERR_ARG, "2^31", 1);
        check(exc_build_index(EXC_SEGLRC, 100, 0.0f, 256, 0, big, off2, 1, 1, &b) == EXC_ERR_ARG, "seg", 1);
        check(exc_build_index(EXC_HS_SEGLRC, 128, 0.0f, 256, 40, big, off2, 1, 1, &b) == EXC_ERR_ARG, "shift", 2);
        check(std::strlen(exc_last_error()) > 0, "message", 2);
    }
    // the synthetic-list generator
    {
        const uint32_t df[] = {0, 1, 100, 70000, 1u << 20};
        std::vector<uint64_t> off(6, 0);
        for (int t = 0; t < 5; ++t) off[t + 1] = off[t] + df[t];
        std::vector<uint32_t> docs(off[5] + 1);
        check(exc_synth_lists(3, 1u << 22, df, 5, 2, docs.data(), off.data()) == EXC_OK, "synth", 0);
        std::vector<uint64_t> short_off = off;
        short_off[3] -= 1;
        check(exc_synth_lists(3, 1u << 22, df, 5, 2, docs.data(), short_off.data()) == EXC_ERR_ARG, "synth off", 0);
        for (int t = 0; t < 5; ++t) {
            check(off[t + 1] - off[t] == df[t], "synth length", 0);
            for (uint64_t i = off[t] + 1; i < off[t + 1]; ++i) check(docs[i] > docs[i - 1], "synth order", 0);
        }
    }
    std::string cmd = std::string("rm -rf ") + dir;
    if (std::system(cmd.c_str()) != 0) std::fprintf(stderr, "could not remove %s\n", dir);
    std::printf("host_asan_check: %s (%d failures)\n", failures ? "FAILED" : "ok", failures);
    return failures ? 1 : 0;
}
