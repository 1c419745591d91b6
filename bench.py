#!/usr/bin/env python3
"""Benchmark: exact k-term intersection queries/s on one MI355X.

Headline (C5): a SegLRC-128 index of 1,000,000 terms over 2^26 docIDs with
df_k = min(2^25, floor(5.8e9 / k^1.1)) -- 25.4 G postings, 40.5 GB resident --
synthesized and encoded on the GPU; one step = one kernelInvoke of a batch of
512 queries of 2-6 terms (drawn proportionally to length, the first queries of
the C5 stream) with index and plan resident in HBM (prologue + probe + scan +
scatter).  Its roofline divides the batch's algorithmic bytes by the probe
kernel's event-timed duration; the CPU baseline runs the oracle on the same
index and a sample of the same batch.  Then the continuous-batching server
takes the C5 stream at 20k queries/s and at doubling offered rates until its
p99 passes 25 ms.

Also measured, on C2 (universe 2^25, 100,000 terms, df_k = min(2^23,
2^25/k); a batch of 1,000 queries of 2-5 terms): SegLRC-128, ParaPFD-128,
LRC-256 and uncompressed lists with each probe kernel's roofline, the plain
single-query GPU path, the hipGraph replay, whole-index decode, file loading,
serving, C1, C3 (hash buckets, compressed and uncompressed) and C4 (learned
windows).  Prints one JSON line.
"""
import argparse
import json
import os
import shutil
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
# k_probe<SegLRC,128> HBM bytes per launch, in KiB, from rocprofv3 --pmc FETCH_SIZE then WRITE_SIZE
# runs of this bench (FETCH x1.99996 and WRITE x1.0, a 1 GiB copy kernel's calibration); they are
# profile figures reported beside this run's, not measured in it
TRAFFIC_C2_KIB = 1464610.0
TRAFFIC_C2_SOURCE = "profiles/k_probe_seglrc128_c2.md (rocprofv3 --pmc on the C2 batch)"
TRAFFIC_C5_KIB = 5952507.0
TRAFFIC_C5_SOURCE = "profiles/k_probe_seglrc128_c5.md (rocprofv3 --pmc on the C5 batch)"
META = {0: 8, 1: 28, 2: 32, 3: 8, 4: 8}  # per-block side data bytes: LRC, SegLRC, HS, ParaPFD, NewPFD
CODEC_NAME = {0: "LRC", 1: "SegLRC", 2: "HS_SegLRC", 3: "ParaPFD", 4: "NewPFD"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


RAW_LINE = 32  # uncompressed lists: docIDs per 128-byte line, the unit an exact search must read


def raw_stages(docs, off):
    """Stage survivors from the raw lists: S_0 = the query's shortest list,
    S_j = S_(j-1) restricted to the (j+1)-th shortest list (stopping at an
    empty set)."""
    def stages(ts):
        out = [docs[int(off[ts[0]]):int(off[ts[0] + 1])]]
        for t in ts[1:]:
            s = out[-1]
            if s.size == 0:
                break
            L = docs[int(off[t]):int(off[t + 1])]
            pos = np.searchsorted(L, s)
            ok = pos < len(L)
            ok[ok] = L[pos[ok]] == s[ok]
            out.append(s[ok])
        return out
    return stages


def algorithmic_bytes(built, docs, off, queries, stages=None, threads=1):
    """Bytes an exact block-decoding algorithm must move per batch: compressed
    words + side data of the shortest list, of every longer-list block that
    holds a probe's target position, and 4 B per result.  The probes of list
    ts[j] are the survivors S_(j-1) of the shorter lists: `stages(ts)` returns
    [S_0, S_1, ...] for a query's terms sorted by length (default: from the
    raw lists docs/off).  Uncompressed lists: the 128-byte lines holding the
    probes' positions (plus, with hash buckets, the two table words of each
    bucket searched).  Returns (algorithmic, decode-everything) bytes."""
    ucnt = built.ind1["m_urlcount"].astype(np.int64)
    wstart = (built.ind1["m_off"] // 4).astype(np.int64)
    wend = np.append(wstart[1:], len(built.ind2))
    codec = built.codec
    raw = codec in (5, 6)
    segnum = None if raw else built.seg_num.astype(np.int64)
    meta = 0 if raw else META[codec]
    segoff = None if raw else built.seg_offset.astype(np.int64)
    if stages is None:
        stages = raw_stages(docs, off)

    def nseg(t):
        if raw:
            return 0
        if codec == 2:
            return int(built.hash_info[built.hash_offset[t] + 1])
        return (int(ucnt[t]) + built.seg - 1) // built.seg

    def list_bytes(t):
        return int(wend[t] - wstart[t]) * 4 + nseg(t) * meta + (12 if codec == 0 else 0)

    def block_words(t, blocks):
        s0, ns = int(segnum[t]), nseg(t)
        o = segoff[s0:s0 + ns]
        ends = np.append(o[1:], wend[t] - wstart[t])
        return int((ends[blocks] - o[blocks]).sum())

    def one(q):
        if not q:
            return 0, 0
        ts = sorted(q, key=lambda t: ucnt[t])  # stable: ascending length, as the planner orders them
        full = sum(list_bytes(t) for t in ts)
        tot = list_bytes(ts[0])
        S = stages(ts)
        for j, t in enumerate(ts[1:]):
            surv = S[j] if j < len(S) else S[-1][:0]
            if surv.size == 0:
                break
            if codec == 0:
                tot += 12
            if raw:
                L = built.ind2[int(wstart[t]):int(wstart[t]) + int(ucnt[t])]
                pos = np.minimum(np.searchsorted(L, surv), max(len(L) - 1, 0)) // RAW_LINE
                lines = np.count_nonzero(np.diff(pos)) + 1
                tot += lines * RAW_LINE * 4
                if codec == 6:
                    h = int(built.hash_offset[t])
                    b = (surv >> built.hash_info[h]).astype(np.int64)
                    tot += 8 * (np.count_nonzero(np.diff(b)) + 1)
                continue
            if codec == 2:
                h = int(built.hash_offset[t])
                b = (surv >> built.hash_info[h]).astype(np.int64)
                b = b[b < int(built.hash_info[h + 1])]
            else:
                s0 = int(segnum[t])
                heads = built.seg_head[s0:s0 + nseg(t)]
                b = np.searchsorted(heads, surv, side="right").astype(np.int64) - 1
                b = b[b >= 0]
            ub = np.unique(b)
            tot += block_words(t, ub) * 4 + len(ub) * meta
        final = S[len(ts) - 1] if len(S) == len(ts) else S[-1][:0]
        return tot + 4 * len(final), full + 4 * len(final)

    if threads > 1:
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(threads) as pool:
            parts = list(pool.map(one, queries))
    else:
        parts = [one(q) for q in queries]
    return sum(p[0] for p in parts), sum(p[1] for p in parts)


def time_batch(ctx, dix, queries, steps, warmup):
    ctx.plan(dix, queries)
    bi = ctx._bi
    for _ in range(warmup):
        ctx.kernelInvoke(bi)
    probe, invoke = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.kernelInvoke(bi)
        p, i = ctx.last_times()
        probe.append(p)
        invoke.append(i)
    dt = time.perf_counter() - t0
    return dt, np.array(probe), np.array(invoke)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


# Host threads of the CPU baseline: the GPU box grants one GPU's job a share
# of 16 host threads, whatever nproc reports, and the baseline stays within it.
CPU_SHARE = 16


def oracle_index(built):
    import oracle
    return oracle.Index(codec=built.codec, seg=built.seg, ucnt=np.ascontiguousarray(built.ind1["m_urlcount"]),
                        uwlist=np.ascontiguousarray(built.ind1["m_off"] // 4), segnum=built.seg_num,
                        words=built.ind2, total_words=len(built.ind2), seg_off=built.seg_offset,
                        seg_head=built.seg_head, seg_median=built.seg_median, lrc=built.lrc_info,
                        fit=built.seg_info, base=built.seg_base, hash_off=built.hash_offset,
                        hash_info=built.hash_info)


def cpu_baseline(built, queries, workload, budget_s=12.0, budget_1t_s=40.0):
    """The oracle's block-skipping decode + intersect on the host cores (the
    cpu_baseline leg): whole-batch repetitions on the box's thread share
    (median of at least 5 when they fit the budget; a batch too slow for that
    is cut to a prefix), per-query latency percentiles from the same runs, and
    one single-thread pass over the batch (or the prefix that fits its
    budget)."""
    import oracle
    ix = oracle_index(built)
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = max(1, min(avail, CPU_SHARE))
    # size the sample: a first pass over a growing prefix until the budget's share is spent
    done, t_used, chunk = 0, 0.0, max(threads, 16)
    while done < len(queries) and t_used < budget_s / 5:
        part = queries[done:done + chunk]
        t0 = time.perf_counter()
        oracle.eval_queries(ix, part, nthreads=threads)
        t_used += time.perf_counter() - t0
        done += len(part)
        chunk = min(chunk * 2, 1024)
    sample = queries[:done]
    reps, qms = [], None
    while len(reps) < 5 or (sum(reps) < budget_s and len(reps) < 50):
        t0 = time.perf_counter()
        _, _, ms = oracle.eval_queries(ix, sample, nthreads=threads, timings=True)
        reps.append(time.perf_counter() - t0)
        qms = ms if qms is None else np.concatenate([qms, ms])
    qps = len(sample) / float(np.median(reps))
    # one thread: the whole sample, or the prefix its budget allows
    per_q1 = float(np.median(reps)) * threads / max(1, len(sample))
    n1 = len(sample) if per_q1 * len(sample) <= budget_1t_s else max(1, int(budget_1t_s / max(per_q1, 1e-9)))
    t0 = time.perf_counter()
    _, _, ms1 = oracle.eval_queries(ix, sample[:n1], nthreads=1, timings=True)
    qps1 = n1 / (time.perf_counter() - t0)
    return {"value": qps, "unit": "queries/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "host_threads_available": avail, "nproc": os.cpu_count(),
            "p50_ms": float(np.percentile(qms, 50)), "p99_ms": float(np.percentile(qms, 99)),
            "repetitions": len(reps), "one_thread_qps": qps1, "one_thread_queries": n1,
            "one_thread_p50_ms": float(np.percentile(ms1, 50)), "one_thread_p99_ms": float(np.percentile(ms1, 99)),
            "sample": "%d of the %d %s queries, oracle block-skipping decode + intersect on %d host threads "
                      "(the box's share for one GPU; %d available, nproc %s, %s), median of %d whole-sample "
                      "repetitions, %.1f s in all, per-query p50/p99 over all of them; 1 thread: %.1f queries/s "
                      "over %d queries"
                      % (len(sample), len(queries), workload, threads, avail, os.cpu_count(), cpu_model(),
                         len(reps), sum(reps), qps1, n1)}


def serving_leg(ea, ctx, dix, queries, batch_res, batch_cum, qps=20000.0, tiles=20, nslots=3, max_batch=256):
    """Poisson arrivals at `qps` of the batch's queries repeated `tiles` times
    through the continuous-batching server; p50/p99 latency, achieved rate and
    whether every repetition returns the batch's results."""
    rng = np.random.default_rng(55)
    allq = queries * tiles
    arrival = np.cumsum(rng.exponential(1.0 / qps, len(allq)))
    cap = len(batch_res) * tiles
    srv = ea.Server(ctx, dix, nslots=nslots, max_batch=max_batch, mem_budget=4 << 30)
    try:
        t0 = time.perf_counter()
        lat, cnt, res = srv.replay(allq, arrival, cap)
        wall = time.perf_counter() - t0
        st = srv.stats()
    finally:
        srv.close()
    counts = np.diff(np.concatenate([[0], batch_cum])).astype(np.uint32)
    same = bool(np.array_equal(cnt, np.tile(counts, tiles)) and np.array_equal(res, np.tile(batch_res, tiles)))
    return {"offered_qps": qps, "queries": len(allq), "achieved_qps": len(allq) / wall,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "slots": nslots, "max_batch": max_batch, "mem_budget": 4 << 30, "batches": st["batches"],
            "overlapped_launches": st["overlapped_launches"], "results_match_batch": same}


def decode_leg(ea, ctx, dix, built, docs, reps=5):
    """Decode every list of the index into HBM (k_decompress); algorithmic
    bytes = compressed words + per-block side data read + 4 B per docID
    written.  The output is compared with the generated lists."""
    terms = np.arange(len(built.ind1), dtype=np.uint32)
    total = int(built.ind1["m_urlcount"].sum())
    out = ctx.device_alloc(total * 4)
    ms = []
    for _ in range(reps + 1):
        ctx.decompress_device(dix, terms, out, total)
        ms.append(ctx.last_times()[0])
    ms = np.array(ms[1:])
    nblocks = int(sum((int(n) + built.seg - 1) // built.seg for n in built.ind1["m_urlcount"]))
    alg = len(built.ind2) * 4 + nblocks * META[built.codec] + total * 4
    same = bool(np.array_equal(ctx.from_device(out, total), docs))
    ctx.device_free(out)
    # ParaPFD, SegLRC and LRC lists: the run decoder (packed run via LDS); the others k_decompress
    kernel = "k_decode_run" if built.codec in (0, 1, 3) else "k_decompress"
    return {"kernel": "%s<%s,%d>" % (kernel, CODEC_NAME[built.codec], built.seg), "docids": total,
            "bits_per_docid": 8.0 * (len(built.ind2) * 4 + nblocks * META[built.codec]) / total,
            "kernel_ms": float(ms.mean()),
            "algorithmic_bytes": int(alg), "achieved_GBps": alg / (ms.mean() * 1e-3) / 1e9,
            "frac_of_peak": alg / (ms.mean() * 1e-3) / 1e9 / HBM_PEAK_GBS, "docids_per_s": total / (ms.mean() * 1e-3),
            "matches_lists": same}


def loader_leg(ea, ctx, built, queries, ref_cum, ref_res, reps=3):
    """The C2 SegLRC-128 index written as the reference's index files, then
    loaded with exc_index_load_files (mapped files, validation, staged upload
    into HBM) with the page cache warm from the write; file bytes per second
    of the whole load, and the loaded index must answer the batch exactly as
    the uploaded one."""
    d = tempfile.mkdtemp(prefix="exc_load_")
    try:
        built.save(d, "c2")
        nbytes = sum(os.path.getsize(os.path.join(d, f)) for f in os.listdir(d))
        ts, lix = [], None
        for _ in range(reps):
            if lix is not None:
                lix.close()
            t0 = time.perf_counter()
            lix = ctx.load_files(ea.SEGLRC, 128, d, "c2")
            ts.append(time.perf_counter() - t0)
        cum, res = ctx.intersect(lix, queries)
        same = bool(np.array_equal(cum, ref_cum) and np.array_equal(res, ref_res))
        lix.close()
    finally:
        shutil.rmtree(d, ignore_errors=True)
    t = float(np.median(ts))
    return {"files": "SegLRC-128 .ind1/.ind2/.segNum/.segOffset/.segHead/.regression_info(_list)/.ind2_distance_base",
            "file_bytes": int(nbytes), "load_s": t, "GBps": nbytes / t / 1e9, "matches_upload": same}


def plain_leg(ea, ctx, dix, qs, ref_cum, ref_res, n=300):
    """The plain single-query GPU path on the first n queries of a config:
    one query per plan / copy / kernels / results sequence, plain binary
    search with one lane per shortest-list element, no work queue.  Its
    results must equal the batched run's (ref_cum / ref_res, cumulative
    counts and results of the same queries).  Leaves the context's options
    at the plain settings: the caller restores its own."""
    ctx.set_tuning(None)
    ctx.set_option(ea.OPT_SEARCH, ea.SEARCH_BINARY)
    ctx.set_option(ea.OPT_PERSISTENT, 0)
    ctx.set_option(ea.OPT_TPE, 1)
    ctx.set_option(ea.OPT_TIMING, 0)
    n = min(n, len(qs))
    lat, same = [], True
    for i in range(n):
        t1 = time.perf_counter()
        _, r = ctx.intersect(dix, [qs[i]])
        lat.append(time.perf_counter() - t1)
        lo = int(ref_cum[i - 1]) if i else 0
        same = same and np.array_equal(r, ref_res[lo:int(ref_cum[i])])
    ctx.set_option(ea.OPT_TIMING, 1)
    ctx.set_option(ea.OPT_TPE, 0)
    return {"queries": n, "qps": n / sum(lat), "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3), "matches_batched": bool(same)}


def learned_leg(ea, W, ctx, dix, df, tune, nq=10_000):
    """C4: 10,000 queries of 2-4 terms (seed 44); the probe kernel with plain
    binary search, the list-fit window and the learned window (PLA-256,
    PLA-Q, then MLP 1-32-1) for lists longer than 65,536; results must be
    identical."""
    qs = W.df_weighted_queries(df, nq, 2, 4, 44)
    coop = tune["coop_max"] if tune else [16] * 6
    out, ref = {"queries": nq, "models_on_lists_longer_than": 65536}, None
    for kind, name in ((ea.MODEL_PLA, "pla"), (ea.MODEL_PLA_Q, "pla_q"), (ea.MODEL_MLP, "mlp")):
        st = ctx.train_models(dix, kind, min_len=65536, epochs=300, samples=4096, lr=1e-3)
        row = {"models": st["nmodels"], "train_ms": st["train_ms"], "mean_err": st["mean_err"], "max_err": st["max_err"]}
        for search, sname in ((0, "binary"), (1, "fit"), (2, "learned")):
            ctx.set_tuning({"coop_max": coop, "search": [search] * 6, "persistent": 0})
            ctx.plan(dix, qs)
            bi = ctx._bi
            ctx.kernelInvoke(bi)
            ms = []
            for _ in range(5):
                ctx.kernelInvoke(bi)
                ms.append(ctx.last_times()[0])
            cum, res = ctx.dtohTransfer(bi)
            if ref is None:
                ref = (cum, res)
            row[sname + "_probe_ms"] = float(np.median(ms))
            row[sname + "_qps"] = nq / (np.median(ms) * 1e-3)
            row[sname + "_mismatches"] = int(0 if (np.array_equal(cum, ref[0]) and np.array_equal(res, ref[1]))
                                             else max(1, int(np.sum(cum != ref[0]))))
        row["learned_gain_vs_binary"] = row["binary_probe_ms"] / row["learned_probe_ms"]
        out[name] = row
    out["plain_single_query"] = plain_leg(ea, ctx, dix, qs, ref[0], ref[1])
    out["batched_learned_vs_plain"] = out["pla"]["learned_qps"] / out["plain_single_query"]["qps"]
    if tune:
        ctx.set_tuning(tune)
    return out


def end_to_end(ea, ctx, dix, queries, reps=5):
    """The batch from host query lists to host results, stage by stage
    (median of reps): the result bound on the host, planning + upload
    (exc_plan), the kernels (exc_kernelInvoke), the results to the host
    (exc_dtohTransfer) into page-locked buffers and into pageable ones; qps
    = queries / the sum of the stages."""
    qterms, qoff = ea.queries_to_flat(queries)
    nq = len(queries)
    cap = ea._result_cap(dix.ucnt, qterms, qoff)
    pin = (ctx.host_empty(nq), ctx.host_empty(cap))
    page = (np.zeros(nq, np.uint32), np.zeros(max(cap, 1), np.uint32))
    st = {k: [] for k in ("bound", "plan_htod", "kernels", "dtoh_pinned", "dtoh_pageable")}
    nres = 0
    for _ in range(reps + 1):
        t0 = time.perf_counter()
        ea._result_cap(dix.ucnt, qterms, qoff)
        t1 = time.perf_counter()
        bi = ctx.plan(dix, (qterms, qoff))
        t2 = time.perf_counter()
        ctx.kernelInvoke(bi)
        t3 = time.perf_counter()
        _, res = ctx.dtohTransfer(bi, cap, out=pin)
        t4 = time.perf_counter()
        ctx.dtohTransfer(bi, cap, out=page)
        t5 = time.perf_counter()
        nres = len(res)
        for k, v in zip(st, (t1 - t0, t2 - t1, t3 - t2, t4 - t3, t5 - t4)):
            st[k].append(v * 1e3)
    ms = {k: float(np.median(v[1:])) for k, v in st.items()}
    pinned = ms["bound"] + ms["plan_htod"] + ms["kernels"] + ms["dtoh_pinned"]
    pageable = ms["bound"] + ms["plan_htod"] + ms["kernels"] + ms["dtoh_pageable"]
    return {"stages_ms": ms, "results": nres, "result_mb": nres * 4 / 1e6,
            "dtoh_pinned_gbs": nres * 4 / (ms["dtoh_pinned"] * 1e6), "qps_pinned": nq / (pinned * 1e-3),
            "qps_pageable": nq / (pageable * 1e-3)}


def clustered_leg(ea, W, ctx, coop=16, nq=10_000):
    """C4s: the learned window where one line fits a list poorly.  Clustered
    lists (shared topic clusters, a staircase docID -> position curve) as
    SegLRC-128 and uncompressed; per layout and model kind (PLA-256, MLP
    1-32-1 on lists longer than 65,536) the 10,000-query batch's probe time
    with plain binary search, the list-fit window (SegLRC's list fit, the
    regression program's window on the uncompressed lists) and the learned
    window; every mode's results must equal binary search's.  The window
    widths say why: the fit's nRangeLeft + nRangeRight against the model's
    2 * (max error + 2), in positions, over the modelled lists."""
    df, docs, off = W.c4s_corpus()
    qs = W.c4s_queries(df, nq)
    out = {"postings": int(len(docs)), "queries": nq, "models_on_lists_longer_than": 65536}
    long_terms = [t for t in range(len(df)) if int(off[t + 1] - off[t]) > 65536]
    for codec, lname in ((ea.SEGLRC, "seglrc128"), (ea.RAW, "uncompressed")):
        b = ea.build_index(docs, off, codec, seg=128)
        dix = ctx.upload(b)
        fit = b.list_info if codec == ea.SEGLRC else b.gu_regression
        fit_w = np.array([int(fit[t]["nRangeLeft"]) + int(fit[t]["nRangeRight"]) for t in long_terms], np.float64)
        lay, ref = {"fit_window_positions_p50": float(np.median(fit_w))}, None
        for kind, name in ((ea.MODEL_PLA_Q, "pla_q"), (ea.MODEL_PLA, "pla"), (ea.MODEL_MLP, "mlp")):
            st = ctx.train_models(dix, kind, min_len=65536, epochs=300, samples=4096, lr=1e-3)
            mw = np.array([2 * (ctx.model_info(dix, t)["err"] + 2) for t in long_terms], np.float64)
            row = {"models": st["nmodels"], "train_ms": st["train_ms"], "model_window_positions_p50": float(np.median(mw))}
            for search, sname in ((ea.SEARCH_BINARY, "binary"), (ea.SEARCH_FIT, "fit"), (ea.SEARCH_LEARNED, "learned")):
                ctx.set_tuning({"coop_max": [coop] * 6, "search": [search] * 6, "persistent": 0})
                ctx.plan(dix, qs)
                bi = ctx._bi
                ctx.kernelInvoke(bi)
                ms = []
                for _ in range(5):
                    ctx.kernelInvoke(bi)
                    ms.append(ctx.last_times()[0])
                cum, res = ctx.dtohTransfer(bi)
                if ref is None:
                    ref = (cum, res)
                row[sname + "_probe_ms"] = float(np.median(ms))
                row[sname + "_matches_binary"] = bool(np.array_equal(cum, ref[0]) and np.array_equal(res, ref[1]))
            row["learned_vs_fit"] = row["fit_probe_ms"] / row["learned_probe_ms"]
            row["learned_vs_binary"] = row["binary_probe_ms"] / row["learned_probe_ms"]
            lay[name] = row
        out[lname] = lay
        log("C4s %s: %s" % (lname, lay))
        dix.close()
        b.close()
    ctx.set_tuning(None)
    return out


def c1_leg(ea, W, ctx, reps=2000):
    """C1, the plain method: one 2-term query (1,000 and 1,000,000 docIDs from
    [0, 2^24), LRC with 256-docID blocks) run `reps` times through the plain
    single-query path (binary search, one lane per element, one launch, host
    planning and PCIe included)."""
    a, b = W.c1_lists(1)
    docs, off = ea.lists_to_flat([a, b])
    built = ea.build_index(docs, off, ea.LRC, seg=256)
    dix = ctx.upload(built)
    want = np.intersect1d(a, b)
    ctx.set_tuning(None)
    ctx.set_option(ea.OPT_SEARCH, ea.SEARCH_BINARY)
    ctx.set_option(ea.OPT_TPE, 1)
    ctx.set_option(ea.OPT_PERSISTENT, 0)
    ctx.set_option(ea.OPT_TIMING, 0)
    lat, ok = [], True
    for _ in range(reps):
        t0 = time.perf_counter()
        cum, res = ctx.intersect(dix, [[0, 1]])
        lat.append(time.perf_counter() - t0)
        ok = ok and np.array_equal(res, want)
    ctx.set_option(ea.OPT_TIMING, 1)
    dix.close()
    lat = np.array(lat) * 1e3
    return {"queries_per_s": reps / (lat.sum() * 1e-3), "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "results": int(len(want)), "exact": bool(ok), "reps": reps}


def ratio_class(df, q):
    """List-length-ratio class of a query (as the library's autotuner forms them)."""
    s = sorted(q, key=lambda t: df[t])
    if len(s) < 2:
        return 5
    r, k = df[s[1]] / max(1, df[s[0]]), 0
    while k < 4 and r >= 10:
        r /= 10
        k += 1
    return k


def routed_leg(ea, ctx, docs, off, hs_keep, seglrc128, qs, sample, rcum, rres, steps):
    """The C3 batch through the library router: candidates are the hash-bucket
    widths of hs_keep and SegLRC at segments of 128 and 256 docIDs; the router
    tunes every candidate, routes each ratio class to its fastest one, keeps
    the routes only if the routed sample beats every candidate alone (its own
    measured runs), then the 10,000-query batch replays as one captured graph
    of all sub-batches."""
    s256 = ea.build_index(docs, off, ea.SEGLRC, seg=256)
    d256 = ctx.upload(s256)
    names = ["hs_" + ("adaptive" if sh == 0 else "shift%d" % sh) for sh in hs_keep] + ["seglrc128", "seglrc256"]
    cands = [v[0] for v in hs_keep.values()] + [seglrc128, d256]
    router = ea.Router(ctx, cands)
    try:
        t = router.autotune(sample, reps=2)
        router.set_option(ea.OPT_GRAPH, 1)
        router.plan(qs)
        router.invoke()
        router.invoke()
        ms = []
        t0 = time.perf_counter()
        for _ in range(steps):
            ms.append(router.invoke())
        dt = time.perf_counter() - t0
        cum, res = router.fetch()
        row = {"candidates": names, "route": {str(k): names[r] for k, r in enumerate(t["route"])},
               "sample_single_ms": dict(zip(names, t["single_ms"])), "sample_routed_ms": t["routed_ms"],
               "persistent": t["persistent"], "graph_step_ms": float(np.median(ms)),
               "qps": len(qs) / (float(np.median(ms)) * 1e-3), "wall_qps": len(qs) * steps / dt,
               "matches_seglrc": bool(np.array_equal(cum, rcum) and np.array_equal(res, rres))}
    finally:
        router.close()
        d256.close()
        s256.close()
    log("C3 routed: %s" % row)
    return row


def c3_leg(ea, W, ctx, docs, off, df, ref_dix, steps, shifts=(13, 12, 14, 0)):
    """C3: the C2 lists in hash buckets on the top docID bits (shift 13 = the top
    12 bits of 2^25; also 12, 14 and the adaptive ~256-per-bucket width), 10,000
    2-term queries with long:short ratios 10^U(1,4).  Per bucket width: search
    knobs autotuned per ratio class, the whole batch as one launch (persistent
    work queue or one wave per unit, as the autotuner chose) replayed from a
    hipGraph, and every class's sub-batch timed.  Then the library's router
    (exc_router_*) takes the four widths and two SegLRC segment sizes as
    candidates, chooses a candidate per ratio class on a 4,000-query sample,
    and the whole 10,000-query batch runs routed: every candidate's sub-batch
    on a stream of its own, one captured hipGraph, timed as one run."""
    qs = W.ratio_queries(df, 10_000, 3)
    sample = W.ratio_queries(df, 4000, 33)
    rcum, rres = ctx.intersect(ref_dix, qs)  # SegLRC result of the same queries
    cls = np.array([ratio_class(df, q) for q in qs])
    classes = [int(k) for k in np.unique(cls)]
    out = {"queries": len(qs), "classes": {str(k): int((cls == k).sum()) for k in classes}, "per_width": {}}
    keep = {}
    for shift in shifts:
        hs = ea.build_index(docs, off, ea.HS_SEGLRC, seg=128, hs_shift=shift)
        dix = ctx.upload(hs)
        ctx.set_option(ea.OPT_GRAPH, 0)
        tune = ctx.autotune(dix, sample, reps=2)
        row = {"device_bytes": dix.device_bytes(), "search": tune["search"], "persistent": tune["persistent"]}
        ctx.set_option(ea.OPT_GRAPH, 1)
        dt, _, inv = time_batch(ctx, dix, qs, steps, 2)
        cum, res = ctx.dtohTransfer(ctx._bi)
        row["graph_qps"] = len(qs) * steps / dt
        row["graph_step_ms"] = float(np.median(inv))
        row["matches_seglrc"] = bool(np.array_equal(cum, rcum) and np.array_equal(res, rres))
        ctx.set_option(ea.OPT_GRAPH, 0)
        ctx.set_option(ea.OPT_PERSISTENT, 0)
        row["class_probe_ms"] = {}
        for k in classes:
            sub = [q for q, c in zip(qs, cls) if c == k]
            _, pr, _ = time_batch(ctx, dix, sub, 5, 1)
            row["class_probe_ms"][str(k)] = float(np.median(pr))
        out["per_width"]["adaptive" if shift == 0 else "shift%d" % shift] = row
        keep[shift] = (dix, hs, row)
        log("C3 shift %d: %s" % (shift, row))
    ctx.set_tuning(None)
    out["fixed_shift13_qps"] = keep[13][2]["graph_qps"]
    out["plain_single_query_shift13"] = plain_leg(ea, ctx, keep[13][0], qs, rcum, rres)
    ctx.set_option(ea.OPT_SEARCH, 1)
    out["routed"] = routed_leg(ea, ctx, docs, off, keep, ref_dix, qs, sample, rcum, rres, steps)
    for dix, hs, _ in keep.values():
        dix.close()
        hs.close()
    # the uncompressed hash-segment layout (GU/cuda_hashSegment_new): the docIDs
    # themselves plus [shift][start_0 .. start_nb] tables, the same queries
    out["uncompressed_hash_segment"] = {}
    for shift in (13, 0):
        rb = ea.build_index(docs, off, ea.RAW_HS, hs_shift=shift)
        dix = ctx.upload(rb)
        ctx.set_option(ea.OPT_GRAPH, 0)
        tune = ctx.autotune(dix, sample, reps=2)
        ctx.set_option(ea.OPT_GRAPH, 1)
        dt, pr, inv = time_batch(ctx, dix, qs, steps, 2)
        cum, res = ctx.dtohTransfer(ctx._bi)
        ctx.set_option(ea.OPT_GRAPH, 0)
        _, pr, _ = time_batch(ctx, dix, qs, steps, 1)  # probe time outside the graph (events)
        alg, _ = algorithmic_bytes(rb, docs, off, qs, threads=CPU_SHARE)
        row = {"device_bytes": dix.device_bytes(), "search": tune["search"], "persistent": tune["persistent"],
               "graph_qps": len(qs) * steps / dt, "graph_step_ms": float(np.median(inv)),
               "matches_seglrc": bool(np.array_equal(cum, rcum) and np.array_equal(res, rres)),
               "roofline": probe_roofline("k_probe<RAW_HS>", alg, pr.mean())}
        out["uncompressed_hash_segment"]["adaptive" if shift == 0 else "shift%d" % shift] = row
        log("C3 uncompressed hash segments shift %d: %s" % (shift, row))
        dix.close()
        rb.close()
        ctx.set_tuning(None)
    ctx.set_option(ea.OPT_SEARCH, 1)
    return out


def gpu_stages(ctx, dix, queries, lengths):
    """Stage survivors of every query from the device itself: S_j is the
    result of the query's j + 1 shortest terms (a batch of all prefixes;
    the batched path's results are checked against host intersections of the
    raw lists in the C5 GPU test).  Returns stages(ts) for algorithmic_bytes."""
    prefixes, key = [], []
    for q in queries:
        ts = tuple(sorted(q, key=lambda t: lengths[t]))
        for j in range(1, len(ts) + 1):
            prefixes.append(list(ts[:j]))
            key.append((ts, j - 1))
    cum, res = ctx.intersect(dix, prefixes)
    cum = np.asarray(cum, np.int64)
    start = np.concatenate([[0], cum[:-1]])
    table = {}
    for (ts, j), b, e in zip(key, start, cum):
        table.setdefault(ts, []).append(res[b:e])
    return lambda ts: table[tuple(ts)]


def c5_headline(ea, W, ctx, steps, warmup, nq=512, cpu=True):
    """C5 batched: the SegLRC-128 index of 1M terms over 2^26 docIDs with
    df_k = min(2^25, floor(C / k^1.1)), C = 5.8e9 (about 40 GB resident),
    synthesized and encoded on the GPU; one step = one kernelInvoke of the
    first `nq` queries of the C5 stream (2-6 terms drawn proportionally to
    length, seed 55) with index and plan resident; knobs autotuned per ratio
    class on another sample (seed 56).  Returns the open index, its list
    offsets and the measurements."""
    U = W.C5_UNIVERSE
    df = W.c5_df()
    t0 = time.perf_counter()
    dix, off = ctx.synth_index_device(W.C5_SEED, U, df, seg=128, batch_postings=1 << 30)
    build_s = time.perf_counter() - t0
    lengths = np.diff(off).astype(np.int64)
    log("C5 index: %d postings, %.1f GB resident, built on the GPU in %.1f s"
        % (int(off[-1]), dix.device_bytes() / 1e9, build_s))
    queries = W.df_weighted_queries_cdf(lengths, nq, 2, 6, 55)
    ctx.set_tuning(None)
    ctx.set_option(ea.OPT_SEARCH, 1)
    ctx.set_option(ea.OPT_TPE, 0)
    tune = ctx.autotune(dix, W.df_weighted_queries_cdf(lengths, 2000, 2, 6, 56), reps=2)
    log("C5 autotune: %s" % tune)
    dt, probe, invoke = time_batch(ctx, dix, queries, steps, warmup)
    bi = ctx._bi
    cum, res = ctx.dtohTransfer(bi)
    cs = ctx.verify(bi, cum, res)
    value = nq * steps / dt
    short = int(sum(int(lengths[q].min()) for q in map(np.asarray, queries)))
    log("C5 batched: %.0f q/s, step %.3f ms, probe %.3f ms, %d results" % (value, 1e3 * dt / steps, probe.mean(), len(res)))
    row = {"queries": nq, "steps": steps, "warmup": warmup, "value": value, "ms_per_step": 1e3 * dt / steps,
           "step_p50_ms": float(np.percentile(invoke, 50)), "step_p99_ms": float(np.percentile(invoke, 99)),
           "probe_ms": float(probe.mean()), "results": int(len(res)), "checksum": [int(x) for x in cs],
           "shortest_list_docids": short, "postings": int(off[-1]), "terms": len(df), "universe": U,
           "C": W.C5_SCALE, "resident_bytes": int(dix.device_bytes()), "build_s_on_gpu": build_s,
           "tune": tune}
    # exact algorithmic bytes: the block layout from the index, stages from prefix intersections
    t0 = time.perf_counter()
    built = ctx.export(dix)
    stages = gpu_stages(ctx, dix, queries, lengths)
    alg, full = algorithmic_bytes(built, None, None, queries, stages=stages, threads=CPU_SHARE)
    row["algorithmic_bytes"] = int(alg)
    row["decode_everything_bytes"] = int(full)
    log("C5 algorithmic bytes %.1f MB per launch (decode-everything %.1f MB), %.1f s"
        % (alg / 1e6, full / 1e6, time.perf_counter() - t0))
    del stages
    if cpu:
        row["cpu_baseline"] = cpu_baseline(built, queries, "C5")
        log("C5 cpu baseline: %s" % row["cpu_baseline"])
    built.close()
    del built
    return dix, off, lengths, queries, tune, row


def c5_serving(ea, W, ctx, dix, lengths, tune, qps=20000.0, seconds=3.0, check=200, p99_bound_ms=25.0):
    """C5 served: 2-6-term queries (seed 55) arriving as a Poisson process
    through the continuous-batching loop (4 slots, 16 GiB admission budget),
    each replay after an untimed half second of the same stream on the same
    server.
    At the configuration's 20k queries/s: docIDs kept on the device (counts
    and latencies to the host; a sample checked alone against the mix) and
    docIDs copied to the host.  Then the offered rate doubles until the loop
    saturates (achieved < 95 % of offered) or p99 passes the bound: the
    highest rate that held is the sustainable rate at that p99."""
    nq = int(qps * seconds)
    queries = W.df_weighted_queries_cdf(lengths, nq, 2, 6, 55)
    rng = np.random.default_rng(55)
    gaps = rng.exponential(1.0, nq)
    ctx.set_tuning(tune)
    out = {"offered_qps": qps, "queries": nq, "slots": 4, "max_batch": 512, "mem_budget": 16 << 30}
    pick = np.random.default_rng(57).choice(nq, check, replace=False)

    def run(rate, counts_only, qs):
        srv = ea.Server(ctx, dix, nslots=4, max_batch=512, mem_budget=16 << 30, counts_only=counts_only)
        try:
            # half a second of the same stream first, untimed: the slots'
            # device buffers and the pinned result buffers grow to the
            # stream's batch sizes (a growth frees and reallocates), so the
            # timed replay sees the loop in its steady state
            nw = min(len(qs), int(rate * 0.5))
            srv.replay(qs[:nw], np.cumsum(gaps[:nw]) / rate + 0.01, 0)
            st0 = srv.stats()
            arrival = np.cumsum(gaps[:len(qs)]) / rate + 0.05
            t0 = time.perf_counter()
            lat, cnt, _ = srv.replay(qs, arrival, 0)
            wall = time.perf_counter() - t0 - 0.05
            st = srv.stats()
        finally:
            srv.close()
        return lat, cnt, {"offered_qps": rate, "achieved_qps": len(qs) / wall,
                          "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                          "mean_results": float(cnt.mean()), "batches": st["batches"] - st0["batches"],
                          "overlapped_launches": st["overlapped_launches"] - st0["overlapped_launches"],
                          "warmup_queries": nw}

    for mode, counts_only in (("counts", True), ("docids_to_host", False)):
        lat, cnt, row = run(qps, counts_only, queries)
        if counts_only:
            same = True
            for i in pick:
                c, _ = ctx.intersect(dix, [queries[i]])
                same = same and int(c[0]) == int(cnt[i])
            row["alone_vs_mix_checked"] = int(check)
            row["alone_vs_mix_equal"] = bool(same)
        out[mode] = row
        log("C5 %s: %s" % (mode, row))
    # offered-load sweep, docIDs on the device
    sweep, best = [], None
    rate = qps
    while rate <= 64 * qps:
        _, _, row = run(rate, True, queries)
        sweep.append(row)
        log("C5 sweep: %s" % row)
        if row["achieved_qps"] < 0.95 * rate or row["p99_ms"] > p99_bound_ms:
            break
        best = row
        rate *= 2
    out["sweep"] = sweep
    out["p99_bound_ms"] = p99_bound_ms
    out["sustainable_qps_at_p99_bound"] = best["achieved_qps"] if best else None
    out["sustainable_point"] = best
    bcum, bres = ctx.intersect(dix, queries[:300])  # batched, for the plain path's check
    out["plain_single_query"] = plain_leg(ea, ctx, dix, queries, bcum, bres)
    ctx.set_option(ea.OPT_SEARCH, 1)
    ctx.set_tuning(None)
    return out


def copy_calibration(mib):
    """A plain device copy of known size, for scaling rocprofv3 FETCH_SIZE /
    WRITE_SIZE readings of the probe kernel."""
    import torch
    n = mib << 18
    a = torch.ones(n, dtype=torch.int32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 4, "launches": 5}), flush=True)


def probe_roofline(kernel, alg, probe_ms):
    """A probe kernel's algorithmic bytes per launch against its mean launch time."""
    ach = alg / (probe_ms * 1e-3) / 1e9
    return {"kernel": kernel, "probe_ms": float(probe_ms), "algorithmic_bytes": int(alg),
            "achieved_GBps": ach, "frac_of_peak": ach / HBM_PEAK_GBS}


def run_c2(args, ea, W, ctx):
    """The C2 batch (1,000 queries of 2-5 terms over the 100k-term Zipf(1.0)
    index of 2^25 docIDs) on every probe codec, plus the legs that reuse its
    corpus (plain path, graph, decode, load, serving, C1, C3, C4)."""
    t0 = time.perf_counter()
    df = W.zipf_df(args.nterms)
    docs, off = ea.synth_lists(df, W.C2_UNIVERSE, 2)
    queries = W.df_weighted_queries(df, args.nq, 2, 5, 22)
    log("corpus: %d terms, %d postings, %.1f s" % (len(df), len(docs), time.perf_counter() - t0))
    t0 = time.perf_counter()
    seg = ea.build_index(docs, off, ea.SEGLRC, seg=128, frac=0.0)
    log("SegLRC-128 built: %.1f MB compressed, %.1f s" % (seg.compressed_bytes() / 1e6, time.perf_counter() - t0))
    t0 = time.perf_counter()
    dix = ctx.upload(seg)
    upload_s = time.perf_counter() - t0
    log("uploaded %.1f MB of index arrays in %.3f s" % (dix.device_bytes() / 1e6, upload_s))
    models = None
    if args.models != "none":
        models = ctx.train_models(dix, ea.MODEL_PLA if args.models == "pla" else ea.MODEL_MLP, min_len=65536,
                                  epochs=300, samples=4096, lr=1e-3)
        log("%s models: %s" % (args.models, models))
    tune = None
    if args.tune_table:
        tune = json.loads(args.tune_table)
    elif args.autotune:
        t0 = time.perf_counter()
        tune = ctx.autotune(dix, W.df_weighted_queries(df, 2000, 2, 5, 23), reps=3)
        log("autotune %.1f s: %s" % (time.perf_counter() - t0, tune))

    def headline_opts():
        ctx.set_option(ea.OPT_SEARCH, args.search)
        ctx.set_option(ea.OPT_PERSISTENT, tune["persistent"] if tune else args.persistent)
        ctx.set_option(ea.OPT_TPE, 0)
        ctx.set_option(ea.OPT_COOP_MAX, args.coop_max)
        if tune:
            ctx.set_tuning(tune)

    headline_opts()
    dt, probe, invoke = time_batch(ctx, dix, queries, args.steps, args.warmup)
    nq = len(queries)
    value = nq * args.steps / dt
    bi = ctx._bi
    cum, res = ctx.dtohTransfer(bi)
    cs = ctx.verify(bi, cum, res)
    log("C2 SegLRC batched: %.0f q/s, step %.3f ms, probe %.3f ms, results %d, checksum %s"
        % (value, 1e3 * dt / args.steps, probe.mean(), len(res), cs))
    t0 = time.perf_counter()
    alg, full = algorithmic_bytes(seg, docs, off, queries, threads=CPU_SHARE)
    log("C2 algorithmic bytes per launch %.1f MB (decode-everything %.1f MB), %.1f s"
        % (alg / 1e6, full / 1e6, time.perf_counter() - t0))
    c2 = {"workload": "C2: 1,000 queries of 2-5 terms, 100k-term Zipf(1.0) index over 2^25 docIDs, "
                      "SegLRC-128 resident in HBM",
          "value": value, "unit": "queries/s", "steps": args.steps, "ms_per_step": 1e3 * dt / args.steps,
          "step_p50_ms": float(np.percentile(invoke, 50)), "step_p99_ms": float(np.percentile(invoke, 99)),
          "queries": nq, "postings": int(len(docs)), "compressed_bytes": int(seg.compressed_bytes()),
          "results": int(len(res)), "checksum": [int(x) for x in cs],
          "knobs": ({"autotuned": True, "coop_max": tune["coop_max"], "search": tune["search"],
                     "persistent": tune["persistent"]} if tune else
                    {"autotuned": False, "search": args.search, "persistent": args.persistent,
                     "coop_max": args.coop_max}),
          "roofline": dict(probe_roofline("k_probe<SegLRC,128>", alg, probe.mean()),
                           decode_everything_bytes=int(full), bound="hbm", peak=HBM_PEAK_GBS, unit="GB/s",
                           traffic=TRAFFIC_C2_KIB * 1024.0, traffic_source=TRAFFIC_C2_SOURCE),
          "roofline_per_codec": {}}
    c2["roofline_per_codec"]["SegLRC-128"] = probe_roofline("k_probe<SegLRC,128>", alg, probe.mean())
    extra = c2
    if not args.no_plain:
        ctx.set_tuning(None)
        ctx.set_option(ea.OPT_SEARCH, ea.SEARCH_BINARY)
        ctx.set_option(ea.OPT_PERSISTENT, 0)
        ctx.set_option(ea.OPT_TPE, 1)
        ctx.set_option(ea.OPT_TIMING, 0)
        nplain = min(nq, 300)
        lat = []
        t0 = time.perf_counter()
        for q in queries[:nplain]:
            t1 = time.perf_counter()
            ctx.intersect(dix, [q])
            lat.append(time.perf_counter() - t1)
        plain_dt = time.perf_counter() - t0
        extra["plain_single_query_qps"] = nplain / plain_dt
        extra["plain_p50_ms"] = float(np.percentile(lat, 50) * 1e3)
        extra["plain_p99_ms"] = float(np.percentile(lat, 99) * 1e3)
        ctx.set_option(ea.OPT_TIMING, 1)
        log("plain single-query: %.0f q/s" % extra["plain_single_query_qps"])
        # batched end to end (plan + htod + kernels + dtoh)
        headline_opts()
        extra["batched_end_to_end"] = end_to_end(ea, ctx, dix, queries)
        extra["batched_end_to_end_qps"] = extra["batched_end_to_end"]["qps_pinned"]
        log("end to end: %s" % extra["batched_end_to_end"])

    # the same batch replayed from a hipGraph (persistent work queue)
    ctx.set_option(ea.OPT_PERSISTENT, 1)
    ctx.set_option(ea.OPT_GRAPH, 1)
    gdt, _, ginv = time_batch(ctx, dix, queries, args.steps, args.warmup)
    gcum, gres = ctx.dtohTransfer(ctx._bi)
    extra["graph_persistent_qps"] = nq * args.steps / gdt
    extra["graph_persistent_step_ms"] = float(np.median(ginv))
    extra["graph_matches"] = bool(np.array_equal(gcum, cum) and np.array_equal(gres, res))
    ctx.set_option(ea.OPT_GRAPH, 0)
    headline_opts()
    log("graph + persistent: %.0f q/s" % extra["graph_persistent_qps"])

    if not args.no_decode:
        extra["decode"] = decode_leg(ea, ctx, dix, seg, docs)
        log("decode: %s" % extra["decode"])

    extra["upload"] = {"device_bytes": int(dix.device_bytes()), "s": upload_s,
                       "GBps": dix.device_bytes() / upload_s / 1e9}
    if not args.no_load:
        extra["load_files"] = loader_leg(ea, ctx, seg, queries, cum, res)
        headline_opts()
        log("load files: %s" % extra["load_files"])

    if not args.no_serve:
        extra["serving"] = serving_leg(ea, ctx, dix, queries, res, cum)
        log("serving: %s" % extra["serving"])

    reps_other = max(3, args.steps // 2)
    if not args.no_pfd:
        pfd = ea.build_index(docs, off, ea.PARAPFD, seg=128, frac=0.1)
        pix = ctx.upload(pfd)
        ctx.set_tuning(None)
        ctx.set_option(ea.OPT_GRAPH, 0)
        if args.autotune and not args.tune_table:
            extra["parapfd_tune"] = ctx.autotune(pix, W.df_weighted_queries(df, 2000, 2, 5, 23), reps=2)
        pdt, pprobe, _ = time_batch(ctx, pix, queries, reps_other, 1)
        extra["parapfd_batched_qps"] = nq * reps_other / pdt
        extra["parapfd_probe_ms"] = float(pprobe.mean())
        pcum, pres = ctx.dtohTransfer(ctx._bi)
        extra["parapfd_matches_seglrc"] = bool(np.array_equal(pcum, cum) and np.array_equal(pres, res))
        palg, _ = algorithmic_bytes(pfd, docs, off, queries, threads=CPU_SHARE)
        c2["roofline_per_codec"]["ParaPFD-128"] = probe_roofline("k_probe<ParaPFD,128>", palg, pprobe.mean())
        log("ParaPFD batched: %.0f q/s, %s" % (extra["parapfd_batched_qps"], c2["roofline_per_codec"]["ParaPFD-128"]))
        if not args.no_decode:
            extra["decode_parapfd"] = decode_leg(ea, ctx, pix, pfd, docs)
            log("decode ParaPFD: %s" % extra["decode_parapfd"])
        pix.close()
        del pfd
        if not args.no_decode:
            # NewPFD (decompression only, as in the reference): 256-docID blocks, FRAC 0.1
            npf = ea.build_index(docs, off, ea.NEWPFD, seg=256, frac=0.1)
            nix = ctx.upload(npf)
            extra["decode_newpfd"] = decode_leg(ea, ctx, nix, npf, docs)
            log("decode NewPFD: %s" % extra["decode_newpfd"])
            nix.close()
            del npf
        headline_opts()

    if not args.no_lrc:
        # LRC (one fit per list, 256-docID residual blocks; the reference's LRC_Intersection)
        lb_ = ea.build_index(docs, off, ea.LRC, seg=256, frac=0.0)
        lix = ctx.upload(lb_)
        ctx.set_tuning(None)
        ctx.set_option(ea.OPT_GRAPH, 0)
        if args.autotune and not args.tune_table:
            extra["lrc_tune"] = ctx.autotune(lix, W.df_weighted_queries(df, 2000, 2, 5, 23), reps=2)
        ldt, lprobe, _ = time_batch(ctx, lix, queries, reps_other, 1)
        lcum, lres = ctx.dtohTransfer(ctx._bi)
        extra["lrc"] = {"compressed_bytes": int(lb_.compressed_bytes()),
                        "batched_qps": nq * reps_other / ldt, "probe_ms": float(lprobe.mean()),
                        "matches_seglrc": bool(np.array_equal(lcum, cum) and np.array_equal(lres, res))}
        lalg, _ = algorithmic_bytes(lb_, docs, off, queries, threads=CPU_SHARE)
        c2["roofline_per_codec"]["LRC-256"] = probe_roofline("k_probe<LRC,256>", lalg, lprobe.mean())
        log("LRC: %s %s" % (extra["lrc"], c2["roofline_per_codec"]["LRC-256"]))
        lix.close()
        lb_.close()
        headline_opts()

    if not args.no_raw:
        # the C2 lists uncompressed (the reference's src/gpu/Uncompressed programs):
        # binary search (mqSearch), the lists' regression window
        # (cuda_runRegressionAlgorithm), their PLA models' window and
        # interpolation search, one wave per unit
        rb = ea.build_index(docs, off, ea.RAW)
        rix = ctx.upload(rb)
        ctx.set_tuning(None)
        extra["uncompressed"] = {"device_bytes": int(rix.device_bytes())}
        extra["uncompressed"]["models"] = ctx.train_models(rix, kind=ea.MODEL_PLA, min_len=65536)
        ralg, _ = algorithmic_bytes(rb, docs, off, queries, threads=CPU_SHARE)
        for mode, name in ((ea.SEARCH_BINARY, "binary"), (ea.SEARCH_FIT, "regression"),
                           (ea.SEARCH_LEARNED, "learned"), (ea.SEARCH_INTERP, "interpolation")):
            ctx.set_option(ea.OPT_SEARCH, mode)
            ctx.set_option(ea.OPT_PERSISTENT, 0)
            rdt, rprobe, _ = time_batch(ctx, rix, queries, reps_other, 1)
            rcum, rres = ctx.dtohTransfer(ctx._bi)
            extra["uncompressed"][name] = {
                "qps": nq * reps_other / rdt, "probe_ms": float(rprobe.mean()),
                "matches_seglrc": bool(np.array_equal(rcum, cum) and np.array_equal(rres, res))}
            c2["roofline_per_codec"]["uncompressed-" + name] = probe_roofline("k_probe<RAW>", ralg, rprobe.mean())
        log("uncompressed: %s" % extra["uncompressed"])
        rix.close()
        rb.close()
        headline_opts()

    if not args.no_c1:
        extra["plain_c1"] = c1_leg(ea, W, ctx)
        log("C1: %s" % extra["plain_c1"])
        headline_opts()

    if not args.no_c3:
        extra["hash_buckets"] = c3_leg(ea, W, ctx, docs, off, df, dix, max(5, args.steps // 2))
        log("C3: %s" % extra["hash_buckets"])
        headline_opts()

    if not args.no_c4:
        extra["learned"] = learned_leg(ea, W, ctx, dix, df, tune)
        log("learned: %s" % extra["learned"])
        extra["learned"]["clustered"] = clustered_leg(ea, W, ctx)
        if tune:
            ctx.set_tuning(tune)

    if not args.no_cpu:
        c2["cpu_baseline"] = cpu_baseline(seg, queries, "C2")
        log("C2 cpu baseline: %s" % c2["cpu_baseline"])
    dix.close()
    seg.close()
    ctx.set_tuning(None)
    return c2


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain single-query leg")
    ap.add_argument("--no-pfd", action="store_true", help="skip the ParaPFD leg")
    ap.add_argument("--nterms", type=int, default=100_000)
    ap.add_argument("--nq", type=int, default=1000)
    ap.add_argument("--c5-nq", type=int, default=512, help="queries in the C5 batch (the headline step)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving legs")
    ap.add_argument("--no-decode", action="store_true", help="skip the whole-index decode leg")
    ap.add_argument("--no-load", action="store_true", help="skip the index-file load leg")
    ap.add_argument("--no-raw", action="store_true", help="skip the uncompressed-list leg")
    ap.add_argument("--no-lrc", action="store_true", help="skip the LRC leg")
    ap.add_argument("--no-c1", action="store_true", help="skip the plain-method single-query leg")
    ap.add_argument("--no-c2", action="store_true", help="skip the C2 legs (and C1, C3, C4, which reuse its corpus)")
    ap.add_argument("--no-c3", action="store_true", help="skip the hash-bucket leg")
    ap.add_argument("--no-c5", action="store_true",
                    help="skip C5 (the headline then falls back to the C2 batch)")
    ap.add_argument("--no-c4", action="store_true", help="skip the learned-window leg")
    ap.add_argument("--models", default="pla", choices=["pla", "mlp", "none"],
                    help="learned position models for lists longer than 65,536 before tuning")
    ap.add_argument("--autotune", type=int, default=1,
                    help="1: pick knobs per ratio class on a separate query sample (seed 23) first")
    ap.add_argument("--tune-table", default=None,
                    help='install this table instead of tuning: {"coop_max": [6], "search": [6], "persistent": p}')
    ap.add_argument("--search", type=int, default=1, help="0 binary, 1 fit window, 2 learned")
    ap.add_argument("--persistent", type=int, default=0)
    ap.add_argument("--coop-max", type=int, default=16)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MiB",
                    help="only run a plain device copy of this size (rocprofv3 counter scaling)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (a batch would shard across "
              "GPUs by replicating the index or by docID range)", file=sys.stderr)
        sys.exit(2)
    if args.no_c2 and args.no_c5:
        print("bench.py: nothing to measure with both --no-c2 and --no-c5", file=sys.stderr)
        sys.exit(2)

    import exclucene_amd as ea
    from exclucene_amd import workloads as W

    ctx = ea.Context(0)
    c2 = None if args.no_c2 else run_c2(args, ea, W, ctx)

    c5 = serving_c5 = None
    if not args.no_c5:
        c5dix, c5off, lengths, c5q, c5tune, c5 = c5_headline(ea, W, ctx, args.steps, args.warmup, nq=args.c5_nq,
                                                            cpu=not args.no_cpu)
        if not args.no_serve:
            serving_c5 = c5_serving(ea, W, ctx, c5dix, lengths, c5tune)
            log("C5 serving: %s" % serving_c5)
        c5dix.close()

    if c5 is not None:
        head = c5
        workload = ("C5: a batch of %d 2-6-term queries (Zipf-weighted by length, seed 55) over a 1M-term "
                    "Zipf(1.1) SegLRC-128 index of %.1f G postings over 2^26 docIDs, %.1f GB resident in HBM"
                    % (c5["queries"], c5["postings"] / 1e9, c5["resident_bytes"] / 1e9))
        roof = dict(probe_roofline("k_probe<SegLRC,128>", c5["algorithmic_bytes"], c5["probe_ms"]),
                    decode_everything_bytes=c5["decode_everything_bytes"],
                    traffic=TRAFFIC_C5_KIB * 1024.0 if TRAFFIC_C5_KIB else None, traffic_source=TRAFFIC_C5_SOURCE)
        cpu = c5.get("cpu_baseline")
        config = {"workload": workload, "queries": c5["queries"], "postings": c5["postings"],
                  "codec": "SegLRC-128", "resident_bytes": c5["resident_bytes"], "knobs": c5["tune"]}
    else:
        head = c2
        roof = c2["roofline"]
        cpu = c2.get("cpu_baseline")
        config = {"workload": c2["workload"], "queries": c2["queries"], "postings": c2["postings"],
                  "codec": "SegLRC-128", "compressed_bytes": c2["compressed_bytes"], "knobs": c2["knobs"]}
    roof = {"bound": "hbm", "achieved": roof["achieved_GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": roof["frac_of_peak"], "traffic": roof.get("traffic"),
            "traffic_source": roof.get("traffic_source"), "kernel": roof["kernel"], "probe_ms": roof["probe_ms"],
            "algorithmic_bytes": roof["algorithmic_bytes"],
            "decode_everything_bytes": roof.get("decode_everything_bytes")}
    out = {
        "metric": "exact k-term intersect queries/s (batched vs plain); % HBM peak; p50/p99 ms",
        "value": head["value"],
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": config,
        "step_p50_ms": head["step_p50_ms"],
        "step_p99_ms": head["step_p99_ms"],
        "results": head["results"],
        "checksum": head["checksum"],
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if c5 is not None:
        out["c5"] = {k: v for k, v in c5.items() if k != "cpu_baseline"}
    if serving_c5 is not None:
        out["serving_c5"] = serving_c5
    if c2 is not None:
        out["c2"] = c2
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
