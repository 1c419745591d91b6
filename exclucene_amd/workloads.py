"""Seeded synthetic workloads (the reference's own datasets -- Baidu bd* and
GOV2 -- are not available; these restate the benchmark configurations):

C1  one 2-term query: 1,000 and 1,000,000 docIDs uniform without replacement
    from [0, 2^24), LRC (one global fit, 256-docID residual blocks).
C2  universe 2^25, 100,000 terms, df_k = min(2^23, floor(2^25 / k)) (Zipf s=1
    capped at 2^25/4); 1,000 queries of 2-5 terms drawn proportionally to df;
    lists encoded as SegLRC-128 and ParaPFD-128 (FRAC 0.1).
C3  the C2 lists split into hash buckets on the top 12 docID bits (shift 13);
    10,000 2-term queries with short:long ratios 10^U(1,4).
C4  the C2 lists with a learned position model on every list longer than
    65,536; 10,000 queries of 2-4 terms.
C5  universe 2^26, 1,000,000 terms, df_k = min(2^25, floor(C / k^1.1)) with
    C = 5.8e9, lists synthesized and SegLRC-encoded on the GPU (about 40 GB
    resident);
    2-6-term queries drawn proportionally to df arriving at 20k per second.
C4s the learned-window workload on clustered lists (no reference counterpart:
    the reference's datasets are crawled documents, whose docIDs bunch by site
    and crawl order): universe 2^25, 10,000 terms, df_k = min(2^21,
    floor(2^23 / k)); 64 shared topic clusters at uniform centres with
    half-widths 2^U(-1,3) * 2^14; each term draws 90 % of its docIDs from 2-10
    of them (Zipf 1.2 weights) and 10 % uniformly, so a long list's docID ->
    position curve is a staircase that one line fits poorly; 10,000 queries of
    2-4 terms drawn proportionally to df (seed 45).
Lists come from exc_synth_lists (splitmix64 per list, Floyd sampling for dense
lists, sort-unique-refill for sparse ones); query streams from numpy PCG64.
"""
import numpy as np

from . import lib as _lib  # noqa: F401  (fail loudly if the library is missing)
from . import synth_lists

C2_UNIVERSE = 1 << 25
C2_TERMS = 100_000
C2_CAP = 1 << 23


def zipf_df(nterms=C2_TERMS, universe=C2_UNIVERSE, cap=C2_CAP, s=1.0, scale=None):
    k = np.arange(1, nterms + 1, dtype=np.float64)
    c = float(universe if scale is None else scale)
    df = np.floor(c / k ** s)
    return np.minimum(df, cap).astype(np.uint32)


C5_UNIVERSE = 1 << 26
C5_TERMS = 1_000_000
C5_SCALE = 5.8e9  # C of df_k = min(2^25, C / k^1.1): about 25.4G postings, 40 GB resident
C5_SEED = 5


def c5_df():
    """Expected list lengths of the C5 index (the device synthesizer draws
    each list as a Bernoulli(df / universe) sample, so lengths vary around df)."""
    return zipf_df(C5_TERMS, universe=C5_UNIVERSE, cap=1 << 25, s=1.1, scale=C5_SCALE)


def c1_lists(seed=1):
    rng = np.random.default_rng(seed)
    a = np.sort(rng.choice(1 << 24, 1000, replace=False)).astype(np.uint32)
    b = np.sort(rng.choice(1 << 24, 1_000_000, replace=False)).astype(np.uint32)
    return [a, b]


def c2_corpus(seed=2, nterms=C2_TERMS, nthreads=0):
    df = zipf_df(nterms)
    docs, off = synth_lists(df, C2_UNIVERSE, seed, nthreads)
    return df, docs, off


def df_weighted_queries(df, nq, tmin, tmax, seed):
    """nq queries; term count uniform in [tmin, tmax]; distinct terms per query,
    drawn with probability proportional to df."""
    rng = np.random.default_rng(seed)
    p = df.astype(np.float64) / float(df.sum())
    out = []
    for _ in range(nq):
        tc = int(rng.integers(tmin, tmax + 1))
        out.append([int(t) for t in rng.choice(len(df), tc, replace=False, p=p)])
    return out


def df_weighted_queries_cdf(df, nq, tmin, tmax, seed):
    """As df_weighted_queries for large vocabularies: terms drawn by inverse
    CDF (with replacement, repeats within a query redrawn)."""
    rng = np.random.default_rng(seed)
    cdf = np.cumsum(df.astype(np.float64))
    cdf /= cdf[-1]
    out = []
    for _ in range(nq):
        tc = int(rng.integers(tmin, tmax + 1))
        q = []
        while len(q) < tc:
            t = int(np.searchsorted(cdf, rng.random(), side="right"))
            t = min(t, len(df) - 1)
            if t not in q:
                q.append(t)
        out.append(q)
    return out


def c2_queries(df, nq=1000, seed=22):
    return df_weighted_queries(df, nq, 2, 5, seed)


def ratio_queries(df, nq=10_000, seed=3, lo=1.0, hi=4.0):
    """2-term queries whose long:short length ratio is 10^U(lo, hi) (df_k ~ C/k,
    so the long term index is the short one divided by the ratio)."""
    rng = np.random.default_rng(seed)
    n = len(df)
    out = []
    for _ in range(nq):
        r = 10.0 ** rng.uniform(lo, hi)
        ks = int(np.exp(rng.uniform(np.log(r), np.log(n))))
        kl = max(1, int(round(ks / r)))
        out.append([ks - 1, kl - 1])
    return out


def flat(queries):
    from . import queries_to_flat
    return queries_to_flat(queries)


C4S_UNIVERSE = 1 << 25
C4S_TERMS = 10_000


def clustered_list(rng, n, centres, halfw, universe=C4S_UNIVERSE):
    """n distinct docIDs, 90 % from a few of the shared clusters, 10 % uniform."""
    if n == 0:
        return np.zeros(0, np.uint32)
    k = int(rng.integers(2, 11))
    pick = rng.choice(len(centres), k, replace=False)
    w = 1.0 / np.arange(1, k + 1) ** 1.2
    w /= w.sum()
    lo = np.maximum(centres[pick] - halfw[pick], 0)
    hi = np.minimum(centres[pick] + halfw[pick], universe)
    ncl = int(n * 0.9)
    c = rng.choice(k, ncl, p=w)
    v = lo[c] + (rng.random(ncl) * (hi - lo)[c]).astype(np.int64)
    v = np.unique(np.concatenate([v, rng.integers(0, universe, n - ncl)]))
    while len(v) < n:  # saturated clusters: the rest uniformly
        v = np.unique(np.concatenate([v, rng.integers(0, universe, (n - len(v)) * 2 + 16)]))
    if len(v) > n:
        v = np.sort(rng.choice(v, n, replace=False))
    return v.astype(np.uint32)


def c4s_corpus(seed=7, nterms=C4S_TERMS):
    """The clustered corpus of C4s: (df, docs, off)."""
    df = zipf_df(nterms, universe=C4S_UNIVERSE, cap=1 << 21, scale=1 << 23)
    rng = np.random.default_rng(seed)
    centres = rng.integers(0, C4S_UNIVERSE, 64)
    halfw = (2.0 ** rng.uniform(-1, 3, 64) * (1 << 14)).astype(np.int64)
    lists = [clustered_list(rng, int(n), centres, halfw) for n in df]
    off = np.zeros(nterms + 1, np.uint64)
    off[1:] = np.cumsum([len(x) for x in lists])
    return df, np.concatenate(lists), off


def c4s_queries(df, nq=10_000, seed=45):
    return df_weighted_queries(df, nq, 2, 4, seed)
