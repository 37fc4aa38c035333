#!/usr/bin/env python3
"""Benchmark of the MI355X hash set-intersection path.

Default workload: the three parts of the headline metric on one GPU, in
ONE JSON line -- set-pair intersections/s (the line's value: 1,000,000 set
pairs, sizes log-uniform in [16, 65536], sorted unique uint32 elements from a
2^24 universe, overlap fraction U[0, 0.5], all pairs in one persistent
size-binned batched launch per step), triangles/s of R-MAT scale 24 (and the
scale-16 graph beside it), and IR p50/p99 latency at fixed QPS over the
10M-document index.

    python bench.py [--steps K] [--warmup W]            # all (default)
    python bench.py --workload pairs                    # the pairs block alone
    python bench.py --workload tc --scale 24            # R-MAT triangle counting
    python bench.py --workload ir                       # IR batch + serving sweep
    python bench.py --workload compare                  # RESET vs the competitor tables

A step = plan (size binning) + the batched intersection kernel over all pairs,
inputs resident in HBM.  `roofline.achieved` = algorithmic bytes of one launch
of the dominant kernel (4(|A|+|B|) + 20 per pair) / its mean duration, timed
with HIP events on the kernel's own stream.  `cpu_baseline` times the CPU port
of the reference algorithm (oracle/) on a bounded sample of the same pairs.
"""
from __future__ import annotations

import argparse
import contextlib
import gc
import json
import os
import sys
import time

HBM_PEAK_GBS = 8000.0
SEED = 20261203


def _cpu_model():
    """The host CPU's model name (first /proc/cpuinfo entry), for the CPU baseline."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


_FLUSH = []
FLUSH_BYTES = 1 << 30


def _flush_caches():
    """Write 1 GiB on the current stream, so that nothing a kernel reads is
    left in the eight 4 MiB L2s or the 256 MiB MALL (the reference flushes
    before timed reps too: baseline/common/utils.cuh:25-41, 60-80)."""
    import torch

    if not _FLUSH:
        _FLUSH.append(torch.empty(FLUSH_BYTES // 4, dtype=torch.int32, device="cuda"))
    _FLUSH[0].fill_(0x5A5A5A5A)


def _cold(run_timed, reps):
    """Median over reps of run_timed() (which enqueues one timed run and
    returns its kernel ms) with the caches flushed before each rep."""
    import numpy as np

    ms = []
    for _ in range(reps):
        _flush_caches()
        ms.append(run_timed())
    return {"kernel_ms_median": float(np.median(ms)), "kernel_ms_min": float(np.min(ms)), "reps": reps,
            "flush_bytes": FLUSH_BYTES}


def _traffic(name):
    """HBM bytes per launch measured with rocprofv3 --pmc (profiles/*.json)."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", f"{name}_traffic.json")
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def _sync_all():
    import torch

    torch.cuda.synchronize()


def bench_pairs(args):
    import numpy as np
    import torch

    from reset_amd import api, inputs

    P = args.pairs
    t0 = time.time()
    batch = inputs.make_pair_batch(P, seed=SEED, lo=args.pair_lo, hi=args.pair_hi)
    _sync_all()
    gen_s = time.time() - t0
    a_off, a, b_off, b, k = batch["a_off"], batch["a"], batch["b_off"], batch["b"], batch["k"]
    n_elems = int(a_off[-1].item()) + int(b_off[-1].item())
    alg_bytes = 4 * n_elems + 20 * P

    opts = api.default_pairs_opts()
    tune = None
    if not args.no_autotune:
        # per-bin (load factor, bucket size, chunk, lanes per pair) by timing
        # candidates on a sample (not timed)
        S = min(P, args.tune_sample)
        t1 = time.time()
        opts = api.autotune_pairs(a_off[:S + 1].contiguous(), a, b_off[:S + 1].contiguous(), b, opts, reps=2)
        _sync_all()
        tune = {"sample_pairs": S, "tune_s": time.time() - t1,
                "bins_alpha_b_chunk_unroll_lanes": api.pairs_opts_table(opts)}
    plan = api.PairsPlan(P, opts)
    sizes = torch.empty(P, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        plan.run_async(a_off, a, b_off, b, sizes)
    _sync_all()
    # timed region: K back-to-back steps, with HIP events on the launch
    # stream around them (the roofline's kernel time: the batched launches
    # and the plan's four bookkeeping kernels, 0.07 ms, back to back)
    _sync_all()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        plan.run_async(a_off, a, b_off, b, sizes)
    e1.record(stream)
    _sync_all()
    wall = time.perf_counter() - t0
    ms_step = wall * 1000.0 / args.steps
    dev_ms = e0.elapsed_time(e1) / args.steps
    plan.status()
    parity = bool(torch.equal(sizes, k))
    # each launch alone: the library's events around the batched launches
    # (after a synchronization, so they include the host's launch gap)
    plan.set_timing(True)
    kms = []
    for _ in range(args.steps):
        plan.run_async(a_off, a, b_off, b, sizes)
        kms.append(plan.kernel_ms())
    def _one():
        plan.run_async(a_off, a, b_off, b, sizes)
        return plan.kernel_ms()

    cold = _cold(_one, max(10, args.steps))
    plan.set_timing(False)
    parity = parity and bool(torch.equal(sizes, k))
    kernel_ms = dev_ms
    achieved = alg_bytes / (kernel_ms / 1000.0) / 1e9

    out = {
        "metric": "set-pair intersections/s (1 GPU)",
        "value": P / (ms_step / 1000.0),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded on device: log-uniform sizes, Feistel-permuted 2^24 universe)",
        "config": {
            "workload": f"batched pairwise intersection: {P} set pairs, |A|,|B| log-uniform in "
                        f"[{args.pair_lo},{args.pair_hi}], sorted unique uint32 from 2^24, overlap U[0,0.5], "
                        "persistent size-binned launches",
            "pairs": P,
            "elements": n_elems,
            "parallelism": "single GPU",
        },
        "parity": {"bit_exact": parity, "check": "sizes == constructed overlap k for all pairs"},
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": "pairs_batched_kernel",
            "kernel_ms": kernel_ms,
            "kernel_ms_isolated": float(np.mean(kms)),
            "alg_bytes_per_launch": alg_bytes,
        },
        "cold_cache": cold,
        "gen_s": gen_s,
        "autotune": tune,
    }
    tr = _traffic("pairs_batched")
    if tr and tr.get("pairs") == P:
        out["roofline"]["traffic"] = tr.get("hbm_bytes_per_launch")
        out["roofline"]["traffic_source"] = tr.get("source")

    if not args.no_plain:
        Np = min(args.plain_sample, P)
        sa, sb = a_off[:Np + 1].contiguous(), b_off[:Np + 1].contiguous()
        api.pairs_intersect_plain(sa, a, sb, b)  # warm-up
        _sync_all()
        t0 = time.perf_counter()
        plain = api.pairs_intersect_plain(sa, a, sb, b)
        _sync_all()
        pt = time.perf_counter() - t0
        out["plain_form"] = {
            "value": Np / pt,
            "unit": "pairs/s",
            "sample_pairs": Np,
            "bit_exact": bool(torch.equal(plain, k[:Np])),
            "batched_speedup": (P / (ms_step / 1000.0)) / (Np / pt),
        }

    if not args.no_cpu:
        import oracle

        S = min(args.cpu_sample, P)
        ha = a_off[:S + 1].cpu().numpy()
        hb = b_off[:S + 1].cpu().numpy()
        A = a[:int(ha[-1])].cpu().numpy()
        B = b[:int(hb[-1])].cpu().numpy()
        threads = oracle.max_threads()
        oracle.pairs_hash(ha, A, hb, B, nthreads=threads)  # warm-up
        # median of 5 runs; the sorted-merge intersection of the same sample
        # is reported beside it as a sanity baseline
        hts, mts = [], []
        for _ in range(5):
            t0 = time.perf_counter()
            ref = oracle.pairs_hash(ha, A, hb, B, nthreads=threads)
            hts.append(time.perf_counter() - t0)
        for _ in range(5):
            t0 = time.perf_counter()
            mref = oracle.pairs_merge(ha, A, hb, B, nthreads=threads)
            mts.append(time.perf_counter() - t0)
        ct = float(np.median(hts))
        out["cpu_baseline"] = {
            "value": S / ct,
            "runs": 5,
            "merge_value": S / float(np.median(mts)),
            "merge_equal": bool(np.array_equal(mref, ref)),
            "unit": "pairs/s",
            "cores": threads,
            "cpu_model": _cpu_model(),
            "kind": "port",
            "sample": f"first {S} of the {P} pairs (same data), CPU port of set_size/intersection.cu "
                      f"(table of the larger set, smaller probes), OpenMP dynamic schedule",
            "bit_exact_vs_gpu": bool(np.array_equal(ref, sizes[:S].cpu().numpy())),
        }
    plan.close()
    return out


def bench_tc(args, scale=None):
    import numpy as np
    import torch

    from reset_amd import api, inputs

    scale = args.scale if scale is None else scale
    t0 = time.time()
    off, cols = inputs.rmat_tc_csr(scale, 16, seed=SEED + 3 if scale == 24 else SEED + scale)
    torch.cuda.empty_cache()
    t = api.HashTables(off, cols, 0.2, 5, 3)
    keys = t.probe_order(1)
    probes, edges = t.tc_work()
    staged, units, lmax = t.tc_staged()
    _sync_all()
    prep_s = time.time() - t0
    out_t = torch.zeros(1, dtype=torch.int64, device="cuda")
    res = {}
    t.set_timing(True)
    for name, layout, pk, plain in (("tiled_hierarchical", 1, keys, False), ("tiled_normal", 0, None, False),
                                    ("plain_hierarchical", 1, keys, True)):
        for _ in range(args.warmup):
            t.tc_count_async(out_t, layout, pk, plain)
        _sync_all()
        st = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        steps = args.steps if not plain else max(1, args.steps // 4)
        e0.record(st)
        for _ in range(steps):
            t.tc_count_async(out_t, layout, pk, plain)
        e1.record(st)
        _sync_all()
        t.status()
        kms = []
        if not plain:  # the count kernels alone, back to back (events inside the library)
            for _ in range(steps):
                t.tc_count_async(out_t, layout, pk, plain)
                kms.append(t.kernel_ms())
        res[name] = (e0.elapsed_time(e1) / steps, int(out_t.item()), float(np.mean(kms)) if kms else None)

    def _one():
        t.tc_count_async(out_t, 1, keys, False)
        return t.kernel_ms()

    cold = _cold(_one, 10)
    counts = {v[1] for v in res.values()} | {int(out_t.item())}
    # the roofline's kernel time: counts back to back, HIP events on the
    # launch stream (the staged and tiled kernels and the counter reset)
    ms = res["tiled_hierarchical"][0]
    kernel_ms = ms
    tri = res["tiled_hierarchical"][1]
    alg_bytes = 8 * probes + 24 * edges
    achieved = alg_bytes / (kernel_ms / 1000.0) / 1e9
    # table build (fill + insert) per layout, the phases the reference times
    # (hashtable.cu:164-177); build bytes 8 E + 4 b BN + 12 n per layout
    d = t.describe()
    build = {}
    for name, mask in (("normal", 1), ("hierarchical", 2)):
        t.rebuild_async(mask)
        _sync_all()
        st = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = max(3, args.steps // 4)
        e0.record(st)
        for _ in range(reps):
            t.rebuild_async(mask)
        e1.record(st)
        _sync_all()
        t.status()
        bms = e0.elapsed_time(e1) / reps
        bbytes = 8 * d.num_elements + 4 * d.bucket_size * d.bucket_num + 12 * d.num_sets
        build[name] = {"ms": bms, "alg_bytes": bbytes, "GBps": bbytes / (bms / 1e3) / 1e9,
                       "frac": bbytes / (bms / 1e3) / 1e9 / HBM_PEAK_GBS,
                       "traffic": _tc_build_traffic(scale, bbytes)}
    # the rebuilt tables count the same
    t.tc_count_async(out_t, 1, keys, False)
    _sync_all()
    counts.add(int(out_t.item()))
    t.set_timing(False)
    out = {
        "metric": "triangles/s (1 GPU)",
        "value": tri / (ms / 1000.0),
        "unit": "triangles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic R-MAT (a=.57 b=.19 c=.19, seeded, labels scrambled)",
        "config": {"workload": f"triangle counting, R-MAT scale {scale} edge factor 16, degree-oriented int32 CSR",
                   "scale": scale, "vertices": int(off.numel() - 1), "oriented_edges": edges,
                   "probes": probes, "parallelism": "single GPU"},
        "lds_staged": {"probes": staged, "share": staged / max(probes, 1), "units": units, "max_table_len": lmax},
        "triangles": tri,
        "parity": {"layouts_agree": len(counts) == 1},
        "pairs_per_s": edges / (ms / 1000.0),
        "variants_ms": {k: v[0] for k, v in res.items()},
        "plain_form": {"value": tri / (res["plain_hierarchical"][0] / 1000.0), "unit": "triangles/s",
                       "ms": res["plain_hierarchical"][0]},
        "cold_cache": cold,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": _tc_traffic(scale, alg_bytes),
                     "kernel": "tc_staged_kernel + tc_tiled_kernel", "kernel_ms": kernel_ms,
                     "kernel_ms_isolated": res["tiled_hierarchical"][2], "alg_bytes_per_launch": alg_bytes},
        "build": build,
        "prep_s": prep_s,
    }
    if not args.no_cpu:
        import oracle

        # the oracle's restatement of TC.cu's hash path: hierarchical tables
        # (alpha 0.2, b 5) built on the CPU first (sequential, reported, not
        # timed), then the probe phase, as the reference times only its probe
        # kernel.  Bounded sample: the edges whose source vertex lies in the
        # first rows of the oriented CSR (labels are scrambled, so a leading
        # row range is a uniform sample); the range grows until ~10 s of CPU
        # probing, or covers the whole graph.  The sorted-merge count of the
        # same rows is the sanity baseline.
        offh = off.cpu().numpy()
        colsh = cols.cpu().numpy()
        n = len(offh) - 1
        threads = oracle.max_threads()
        t0 = time.perf_counter()
        ctab = oracle.build_tables(offh, colsh, 0.2, 5, 1)
        build_s = time.perf_counter() - t0
        rows, ct, tri_cpu = max(1, n // 1024), 0.0, 0
        while True:
            t0 = time.perf_counter()
            tri_cpu = oracle.tc_hash_rows(offh, colsh, ctab, 0, rows, nthreads=threads)
            ct = time.perf_counter() - t0
            if ct >= args.cpu_seconds or rows >= n:
                break
            rows = min(n, rows * max(2, int(args.cpu_seconds / max(ct, 1e-3))))
        t0 = time.perf_counter()
        tri_merge = oracle.tc_merge_rows(offh, colsh, 0, rows, nthreads=threads)
        mt = time.perf_counter() - t0
        del ctab
        # the GPU count of the same rows: the plain kernel over the sample's
        # edges is the reference form; here the whole-graph count is compared
        # when the sample is the whole graph, else the sample's CPU hash and
        # merge counts must agree
        full = rows >= n
        out["cpu_baseline"] = {"value": tri_cpu / ct, "unit": "triangles/s", "cores": threads,
                               "cpu_model": _cpu_model(), "kind": "port",
                               "sample": f"triangles with source vertex in rows [0, {rows}) of {n} "
                                         f"({int(offh[rows])} of {int(offh[-1])} oriented edges): probe phase "
                                         "of the CPU restatement of TC.cu (hierarchical tables, alpha 0.2, b 5, "
                                         "smaller list probes, chunk-4 dynamic OpenMP schedule)",
                               "sample_triangles": tri_cpu, "seconds": ct, "table_build_s": build_s,
                               "merge_value": tri_merge / mt, "merge_equal": tri_merge == tri_cpu}
        if full:
            out["cpu_baseline"]["bit_exact_vs_gpu"] = tri_cpu == tri
    t.close()
    return out


def _tc_traffic(scale, alg_bytes):
    """Counter-measured HBM bytes per count when a profile of this workload
    exists (same algorithmic bytes, i.e. the same graph)."""
    tr = _traffic(f"tc_rmat{scale}")
    if tr and tr.get("alg_bytes_per_launch") == alg_bytes:
        return tr["hbm_bytes_per_launch"]
    return None


def _tc_build_traffic(scale, alg_bytes):
    """Counter-measured HBM bytes of one layout's build, when profiled."""
    tr = _traffic(f"tc_rmat{scale}")
    b = tr.get("build", {}) if tr else {}
    if b.get("alg_bytes_per_layout") == alg_bytes:
        return b.get("hbm_bytes_per_layout")
    return None


def bench_sss(args):
    import numpy as np
    import torch

    from reset_amd import api, inputs

    t0 = time.time()
    off, cols = inputs.zipf_set_collection(args.sets, 1_000_000, seed=SEED + 4)
    planted = None
    if args.sss_plant > 0:
        off, cols, planted = inputs.plant_near_duplicates(off, cols, args.sss_plant, seed=SEED + 10)
    _sync_all()
    gen_s = time.time() - t0
    for _ in range(args.warmup):
        api.sss_selfjoin(off, cols, 4, 5)
    _sync_all()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = api.sss_selfjoin(off, cols, 4, 5)
    _sync_all()
    ms = (time.perf_counter() - t0) * 1000.0 / args.steps
    pairs = res["pairs"].cpu().numpy()
    offh, colsh = off.cpu().numpy(), cols.cpu().numpy()
    ok = True
    if planted is not None:  # completeness: every planted near-duplicate pair is reported
        ok_planted = bool(np.isin(planted, pairs).all())
    for key in pairs[:1000].tolist():  # every reported pair meets J >= 4/5 exactly
        i, j = key >> 32, key & 0xFFFFFFFF
        a, b = colsh[offh[i]:offh[i + 1]], colsh[offh[j]:offh[j + 1]]
        r = len(np.intersect1d(a, b, assume_unique=True))
        ok &= 5 * r >= 4 * (len(a) + len(b) - r)
    vms = res["ms"][2]
    # per candidate 8 min(|A|, |B|) + 24 bytes: the smaller set probed (key and
    # home slot) plus ids and offsets; verify_probes is the sum of the mins
    alg = 8 * res["verify_probes"] + 24 * res["candidates"]
    out = {
        "metric": "set-pair intersections/s (1 GPU)",
        "value": res["candidates"] / (vms / 1000.0) if vms > 0 else 0.0,
        "unit": "candidate pairs verified/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic Zipf sets (seeded on device)"
                + (f", last {args.sss_plant} sets replaced by one-token-perturbed copies of earlier sets"
                   if args.sss_plant else ""),
        "config": {"workload": f"set-similarity self-join J>=0.8, {args.sets} sets, sizes Zipf(1.2) in [10,2000], "
                               "tokens Zipf(1.0) over 1M, prefix filter + batched hash verification",
                   "sets": args.sets, "elements": int(offh[-1]), "parallelism": "single GPU"},
        "result_pairs": int(len(pairs)), "candidates": res["candidates"],
        "verify_elements": res["verify_elements"], "verify_probes": res["verify_probes"],
        "stage_ms": {"token_order": res["ms"][0], "candidates": res["ms"][1], "verify": res["ms"][2]},
        "parity": {"exact_jaccard_of_reported_pairs": bool(ok),
                   "planted_pairs_reported": ok_planted if planted is not None else None},
        "roofline": {"bound": "hbm", "achieved": alg / (vms / 1000.0) / 1e9 if vms > 0 else 0.0,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "traffic": None,
                     "kernel": "verify stage (group_verify_kernel over candidates + exact rule + compaction)",
                     "alg_bytes_per_launch": alg},
        "gen_s": gen_s,
    }
    out["roofline"]["frac"] = out["roofline"]["achieved"] / HBM_PEAK_GBS
    tr = _traffic("sss_selfjoin")
    if tr and tr.get("candidates") == res["candidates"]:
        out["roofline"]["traffic"] = tr["hbm_bytes_per_launch"]
        out["roofline"]["traffic_GBps"] = tr["hbm_bytes_per_launch"] / (vms / 1e3) / 1e9
    out["roofline"]["stream_bytes"] = 4 * res["verify_elements"] + 20 * res["candidates"]
    if not args.no_cpu:
        import oracle

        S = min(args.cpu_sample, args.sets)
        o = offh[:S + 1]
        c = colsh[:o[-1]]
        threads = oracle.max_threads()
        t0 = time.perf_counter()
        ref, nc = oracle.sss_prefix_join(o, c, 4, 5, nthreads=threads)
        ct = time.perf_counter() - t0
        got = api.sss_selfjoin(off[:S + 1].contiguous(), cols, 4, 5)
        out["cpu_baseline"] = {"value": nc / ct, "unit": "candidate pairs verified/s (end to end)",
                               "cores": threads, "cpu_model": _cpu_model(), "kind": "port",
                               "sample": f"self-join of the first {S} sets (prefix filter + merge verification)",
                               "bit_exact_vs_gpu": bool(np.array_equal(ref, got["pairs"].cpu().numpy()))
                               and nc == got["candidates"]}
    return out


def _ir_probe_work(t, q, qoff, layout):
    """Sum over queries of |S_0| + ... + |S_{k-2}|, the keys each filter step
    probes (S_0 = term0's list, S_j = survivors of the first j + 1 terms): the
    result counts of the queries cut to their first j + 1 terms, j = 0 .. 6."""
    import numpy as np
    import torch

    k = np.diff(qoff)
    total = 0
    for j in range(int(k.max()) - 1):
        sel = np.flatnonzero(k >= j + 2)  # queries whose step j + 1 probes S_j
        if len(sel) == 0:
            break
        cut = [q[qoff[i]:qoff[i] + j + 1] for i in sel]
        cq = np.concatenate(cut).astype(np.int32)
        co = np.arange(0, (j + 1) * len(sel) + 1, j + 1, dtype=np.int32)
        _, cnt, _ = t.ir_query(torch.from_numpy(cq).cuda(), torch.from_numpy(co).cuda(), layout)
        total += int(cnt.sum().item())
    return total


def _ir_hyper(args, off):
    """(load factor, bucket size) of the IR tables: the reference solver's
    least-cost pair (rs_hyper_solve) for the IR kernels' probe width -- one
    lane per key -- and 128-byte transactions, at the index's mean posting
    length; --ir-alpha / --ir-bucket override it."""
    from reset_amd import api

    n, E = off.numel() - 1, int(off[-1].item())
    alpha, b, _ = api.hyper_solve(max(1, E // max(1, n)), W=1, S_trans=128)
    if args.ir_alpha is not None:
        alpha = args.ir_alpha
    if args.ir_bucket is not None:
        b = args.ir_bucket
    return alpha, b


def bench_ir_batch(args):
    """All IR queries of the 10M-document index in one batch, hierarchical
    tables, K times: the roofline of the query kernels alone (the command the
    IR counter profiles run)."""
    import numpy as np
    import torch

    from reset_amd import api, inputs

    off, docs = inputs.ir_index(seed=SEED + 5)
    alpha, b = _ir_hyper(args, off)
    t = api.HashTables(off, docs, alpha, b, 2)
    q, qoff = inputs.ir_queries(off, args.queries, seed=SEED + 6)
    qd, qod = torch.from_numpy(q).cuda(), torch.from_numpy(qoff).cuda()
    for _ in range(args.warmup):
        t.ir_query(qd, qod, 1)
    t.set_timing(True)
    kms = []
    for _ in range(args.steps):
        _, cnt, roff = t.ir_query(qd, qod, 1)
        kms.append(t.kernel_ms())
    t.set_timing(False)
    probes = _ir_probe_work(t, q, qoff, 1)
    n0 = int(roff[-1].item())
    alg = 8 * n0 + 8 * probes
    ms = float(np.mean(kms))
    t.close()
    return {"metric": "IR batch of all queries: query kernels (1 GPU)", "value": ms, "unit": "ms",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
            "data": "synthetic Zipf(1.1) inverted index (seeded on device)",
            "config": {"workload": f"{len(qoff) - 1} conjunctive queries of 2-8 terms over 10M docs / 1M terms, "
                                   "one batch, hierarchical tables", "layout": "hierarchical",
                       "load_factor": alpha, "bucket_size": b},
            "results": int(cnt.sum().item()),
            "roofline": {"bound": "hbm", "achieved": alg / (ms / 1e3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": alg / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, "kernel_ms": ms,
                         "alg_bytes_per_launch": alg, "term0_entries": n0, "probes": probes}}


@contextlib.contextmanager
def _no_gc():
    """The arrival loop runs without Python's cyclic collector: a full pass
    over the results held by then (millions of ints) pauses the submitting
    thread for ~0.1 s, and the requests due meanwhile would arrive as one
    burst."""
    gc.collect()
    gc.disable()
    try:
        yield
    finally:
        gc.enable()


def _serve_at(t, queries, qps, n, args, seed, layout):
    """Poisson arrivals at `qps` of the first n queries through the serving
    loop -> (latencies ms, span ms, results by id, stats)."""
    import numpy as np

    from reset_amd import serving

    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / qps, n)
    srv = serving.Server(t, layout=layout, budget_bytes=args.ir_budget_mb << 20, max_batch=args.ir_batch,
                         max_inflight=4)
    with _no_gc():
        start = time.perf_counter()
        due = start + np.cumsum(gaps)
        for i in range(n):
            d = due[i] - time.perf_counter()
            if d > 0:
                time.sleep(d)
            srv.submit(i, queries[i])
        srv.drain()
    comps = srv.poll_all()
    lat = np.array([c.done_ms - c.submit_ms for c in comps])
    span_ms = max(c.done_ms for c in comps) - min(c.submit_ms for c in comps)
    served = {c.query_id: srv.result(c.query_id) for c in comps}
    stats = srv.stats()
    srv.close()
    return lat, span_ms, served, stats


def bench_ir(args):
    import numpy as np
    import torch

    from reset_amd import api, inputs

    t0 = time.time()
    off, docs = inputs.ir_index(seed=SEED + 5)
    alpha, b = _ir_hyper(args, off)
    t = api.HashTables(off, docs, alpha, b, 3)
    q, qoff = inputs.ir_queries(off, args.queries, seed=SEED + 6)
    _sync_all()
    prep_s = time.time() - t0
    Q = len(qoff) - 1
    queries = [q[qoff[i]:qoff[i + 1]] for i in range(Q)]
    qd, qod = torch.from_numpy(q).cuda(), torch.from_numpy(qoff).cuda()
    # batch of all queries: hierarchical tables (RESET's layout: the ascending
    # term0 list probes each table's buckets in order), normal beside it
    layout = 1
    t.set_timing(True)
    batch = {}
    for name, lay in (("hierarchical", 1), ("normal", 0)):
        t.ir_query(qd, qod, lay)
        _sync_all()
        kms, walls = [], []
        for _ in range(max(3, args.steps // 2)):
            w0 = time.perf_counter()
            res, cnt, roff = t.ir_query(qd, qod, lay)
            walls.append((time.perf_counter() - w0) * 1e3)
            kms.append(t.kernel_ms())
        batch[name] = {"kernel_ms": float(np.mean(kms)), "call_ms": float(np.median(walls)),
                       "results": int(cnt.sum().item())}
        if lay == layout:
            resh, cnth, roffh = res.cpu().numpy(), cnt.cpu().numpy(), roff.cpu().numpy()
    t.set_timing(False)
    # the same batch over tables with the reference apps' defaults (alpha
    # 0.2, b 5; hierarchical): same results, for comparison
    ref_hp = None
    if (alpha, b) != (0.2, 5):
        tr5 = api.HashTables(off, docs, 0.2, 5, 2)
        tr5.ir_query(qd, qod, layout)
        tr5.set_timing(True)
        kms = []
        for _ in range(3):
            res5, cnt5, _ = tr5.ir_query(qd, qod, layout)
            kms.append(tr5.kernel_ms())
        same5 = bool(np.array_equal(cnt5.cpu().numpy(), cnth) and np.array_equal(res5.cpu().numpy(), resh))
        tr5.close()
        del res5, cnt5
        ref_hp = {"load_factor": 0.2, "bucket_size": 5, "kernel_ms": float(np.mean(kms)), "same_results": same5}
    n0 = int(roffh[-1])
    probes = _ir_probe_work(t, q, qoff, layout)
    alg_bytes = 8 * n0 + 8 * probes
    kernel_ms = batch["hierarchical"]["kernel_ms"]
    achieved = alg_bytes / (kernel_ms / 1e3) / 1e9
    # serving at fixed Poisson arrival rates
    sweep = {}
    served_ok = True
    for qps in args.qps_sweep:
        n = min(Q, max(2000, int(qps * args.serve_seconds)))
        lat, span_ms, served, stats = _serve_at(t, queries, qps, n, args, SEED + 7, layout)
        ok = all(served[i] == resh[roffh[i]:roffh[i] + cnth[i]].tolist() for i in range(n))
        served_ok &= ok
        sweep[str(int(qps))] = {"queries": n, "p50_ms": float(np.percentile(lat, 50)),
                                "p90_ms": float(np.percentile(lat, 90)), "p99_ms": float(np.percentile(lat, 99)),
                                "max_ms": float(lat.max()), "throughput_qps": n / (span_ms / 1000.0),
                                "launches": stats["launches"], "peak_bytes": stats["peak_bytes"],
                                "served_equals_batch": bool(ok)}
    head = sweep.get(str(int(args.qps)), next(iter(sweep.values())))
    out = {
        "metric": "IR query latency p50/p99 at fixed QPS (1 GPU)",
        "value": head["p99_ms"], "unit": "ms (p99)", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": head["p50_ms"], "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic Zipf(1.1) inverted index (seeded on device)",
        "config": {"workload": "IR serving: 10M docs, 1M terms, posting lengths Zipf(1.1) capped at 2M, "
                               f"{Q} conjunctive queries of 2-8 terms, Poisson arrivals",
                   "postings": int(off[-1].item()), "queries": Q, "qps": args.qps, "layout": "hierarchical",
                   "load_factor": alpha, "bucket_size": b,
                   "hyperparameters": "rs_hyper_solve at probe width 1 (one lane per key), 128-byte transactions",
                   "parallelism": "single GPU"},
        "latency_ms": {"p50": head["p50_ms"], "p99": head["p99_ms"], "max": head["max_ms"]},
        "qps_sweep": sweep,
        "batch_all_queries": batch,
        "batch_reference_hyperparameters": ref_hp,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "ir_segment_kernel with its count, scan and copy kernels (all queries in one batch)",
                     "kernel_ms": kernel_ms, "alg_bytes_per_launch": alg_bytes,
                     "term0_entries": n0, "probes": probes},
        "parity": {"served_equals_batch": bool(served_ok)},
        "prep_s": prep_s,
    }
    tr = _traffic("ir_batch")
    if tr and tr.get("alg_bytes_per_launch") == alg_bytes and \
            (tr.get("load_factor"), tr.get("bucket_size")) == (alpha, b):
        out["roofline"]["traffic"] = tr["hbm_bytes_per_launch"]
    if not args.no_cpu:
        import oracle

        # the oracle's restatement of ir.cu's hash path (tables built on the
        # CPU first, not timed), on a leading sample of the queries grown to
        # ~10 s; the chained sorted merge of ALL queries checks the GPU batch
        offh, docsh = off.cpu().numpy(), docs.cpu().numpy()
        threads = oracle.max_threads()
        ctab = oracle.build_tables(offh, docsh, 0.2, 5, 1)
        S, ct = min(Q, 500), 0.0
        while True:
            t0 = time.perf_counter()
            o_res, o_cnt, o_roff = oracle.ir_hash(offh, docsh, ctab, q[:qoff[S]], qoff[:S + 1], nthreads=threads)
            ct = time.perf_counter() - t0
            if ct >= args.cpu_seconds or S >= Q:
                break
            S = min(Q, S * max(2, int(args.cpu_seconds / max(ct, 1e-3))))
        del ctab
        t0 = time.perf_counter()
        m_res, m_cnt, m_roff = oracle.ir_merge(offh, docsh, q, qoff, nthreads=threads)
        mt = time.perf_counter() - t0
        same = np.array_equal(m_cnt, cnth)
        if same:
            keep = np.repeat(np.arange(Q), m_cnt)
            pos = np.repeat(m_roff[:-1], m_cnt) + (np.arange(int(m_cnt.sum())) - np.repeat(np.cumsum(m_cnt) - m_cnt, m_cnt))
            same = bool(np.array_equal(m_res[pos], resh[pos])) and len(keep) == int(cnth.sum())
        out["cpu_baseline"] = {"value": S / ct, "unit": "queries/s", "cores": threads, "cpu_model": _cpu_model(),
                               "kind": "port", "sample": f"the first {S} of the {Q} queries: the CPU restatement "
                               "of ir.cu (hierarchical tables alpha 0.2 b 5, survivors filtered term by term)",
                               "seconds": ct, "merge_value": Q / mt,
                               "bit_exact_vs_gpu": bool(same),
                               "check": "chained sorted merge of all queries == GPU batch (hierarchical) results"}
    if args.sss_lookups > 0:
        out["mixed_serving"] = _mixed_serving(args, t, queries, resh, cnth, roffh)
    t.close()
    return out


def _mixed_serving(args, t, queries, resh, cnth, roffh):
    """IR queries and similarity lookups (J >= 4/5 against a resident Zipf
    set collection) arriving interleaved on one serving loop."""
    import numpy as np
    import torch

    from reset_amd import api, inputs, serving

    soff, scols = inputs.zipf_set_collection(args.sss_sets, 1_000_000, seed=SEED + 8)
    sidx = api.SSSIndex(soff, scols, 4, 5)
    soff_h, scols_h = soff.cpu().numpy(), scols.cpu().numpy()
    rng = np.random.default_rng(SEED + 9)
    n_s = args.sss_lookups
    lookups = []
    for j, i in enumerate(rng.choice(len(soff_h) - 1, n_s).tolist()):
        r = scols_h[soff_h[i]:soff_h[i + 1]].copy()
        if j % 2 and len(r) > 5:
            r[int(rng.integers(0, len(r)))] = 3_000_000 + j  # a token outside the collection
        lookups.append(np.unique(r).astype(np.int32))
    lq = np.zeros(n_s + 1, np.int64)
    np.cumsum([len(r) for r in lookups], out=lq[1:])
    got = sidx.lookup(torch.from_numpy(lq.astype(np.int32)).cuda(), torch.from_numpy(np.concatenate(lookups)).cuda())
    keys = got["pairs"].cpu().numpy()
    batch_sss = [[] for _ in range(n_s)]
    for k in keys.tolist():
        batch_sss[k >> 32].append(k & 0xFFFFFFFF)
    n_ir = min(len(queries), args.queries)
    rate = args.qps + args.sss_qps
    gaps = rng.exponential(1.0 / rate, n_ir + n_s)
    kinds = rng.permutation(np.r_[np.zeros(n_ir, np.int8), np.ones(n_s, np.int8)])
    srv = serving.Server(t, layout=1, budget_bytes=args.ir_budget_mb << 20, max_batch=args.ir_batch,
                         max_inflight=args.mixed_inflight, sss_index=sidx)
    with _no_gc():
        start = time.perf_counter()
        due = start + np.cumsum(gaps)
        ii = si = 0
        for k in range(n_ir + n_s):
            d = due[k] - time.perf_counter()
            if d > 0:
                time.sleep(d)
            if kinds[k] == 0:
                srv.submit_ir(ii, queries[ii])
                ii += 1
            else:
                srv.submit_sss(1 << 40 | si, lookups[si])
                si += 1
        srv.drain()
    comps = srv.poll_all()
    lat = {0: [], 1: []}
    ok_ir = ok_s = True
    for c in comps:
        lat[c.kind].append(c.done_ms - c.submit_ms)
        r = srv.result(c.query_id)
        if c.kind == 1:
            ok_s &= r == batch_sss[c.query_id & ((1 << 40) - 1)]
        else:
            i = c.query_id
            ok_ir &= r == resh[roffh[i]:roffh[i] + cnth[i]].tolist()
    stats = srv.stats()
    srv.close()
    sidx.close()
    pct = lambda v, p: float(np.percentile(np.array(v), p)) if v else None
    return {"ir_queries": n_ir, "sss_lookups": n_s, "arrival_qps": rate, "ir_streams": args.mixed_inflight,
            "sss_collection": {"sets": int(len(soff_h) - 1), "threshold": "4/5"},
            "ir_latency_ms": {"p50": pct(lat[0], 50), "p99": pct(lat[0], 99)},
            "sss_latency_ms": {"p50": pct(lat[1], 50), "p99": pct(lat[1], 99)},
            "sss_matches": int(len(keys)), "server": stats,
            "parity": {"ir_served_equals_batch": bool(ok_ir), "sss_served_equals_batch": bool(ok_s)}}


def _timed(fn, steps, warmup):
    """Average ms of fn() over `steps` calls after `warmup` calls (events on
    the current stream; fn enqueues or synchronizes itself)."""
    import torch

    for _ in range(warmup):
        fn()
    _sync_all()
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(steps):
        fn()
    e1.record(st)
    _sync_all()
    return e0.elapsed_time(e1) / steps


def bench_compare(args):
    """RESET's tables against the baseline directory's competitor tables
    (linear probing, cuckoo, hopscotch, roaring) on the reference's three
    comparisons: TC over a degree-oriented R-MAT CSR, SSS over the edges of a
    symmetric R-MAT CSR at threshold 1/4, and IR conjunctive queries
    (baseline/{tc,sss,ir}/main.cu).  Not the headline metric: a comparison
    line."""
    import torch

    from reset_amd import api, inputs

    kinds = ["linear", "cuckoo", "hopscotch", "roaring"]
    steps, warm = max(1, args.steps // 2), 1
    # ---- TC
    off, cols = inputs.rmat_tc_csr(args.scale, 16, seed=SEED + 3)
    E = int(cols.numel())
    t = api.HashTables(off, cols, 0.2, 5, 3)
    keys = t.probe_order(1)
    out_t = torch.zeros(1, dtype=torch.int64, device="cuda")
    tc = {}
    for name, layout, pk, plain in (("reset_tiled_hierarchical", 1, keys, False),
                                    ("reset_plain_hierarchical", 1, keys, True),
                                    ("reset_plain_normal", 0, None, True)):
        ms = _timed(lambda: t.tc_count_async(out_t, layout, pk, plain), steps, warm)
        tc[name] = {"count_ms": ms, "triangles": int(out_t.item())}
    t.close()
    for k in kinds:
        _sync_all()
        t0 = time.perf_counter()
        c = api.CompetitorTables(off, cols, k, 0.2)
        _sync_all()
        build_ms = (time.perf_counter() - t0) * 1e3
        res = {}
        ms = _timed(lambda: res.__setitem__("n", c.tc_count()), steps, warm)
        tc[k] = {"count_ms": ms, "build_ms": build_ms, "triangles": res["n"], "slots": c.slots,
                 "lost_keys": c.failed}
        c.close()
    # ---- SSS (reference form: candidates = the symmetric CSR's edges)
    lo, hi, n = inputs.rmat_undirected(args.scale - 2, 16, seed=SEED + 4)
    soff, scols = inputs.symmetric_csr(lo, hi, n)
    rows = api.expand_rows(soff)
    P = int(scols.numel())
    t = api.HashTables(soff, scols, 0.2, 5, 3)
    skeys = t.probe_order(1)
    sss = {}
    res = {}
    ms = _timed(lambda: res.__setitem__("v", t.sss_verify(rows, scols, 1, 4)), steps, warm)
    flags0, sizes0 = res["v"]
    sss["reset_normal"] = {"verify_ms": ms, "similar": int(flags0.sum().item())}
    # RESET's method: hierarchical tables, probe keys read in table order
    ms = _timed(lambda: res.__setitem__("v", t.sss_verify(rows, scols, 1, 4, layout=1, probe_keys=skeys)),
                steps, warm)
    sss["reset_hierarchical"] = {"verify_ms": ms, "similar": int(res["v"][0].sum().item()),
                                 "sizes_equal_reset": bool(torch.equal(res["v"][1], sizes0))}
    t.close()
    for k in kinds:
        c = api.CompetitorTables(soff, scols, k, 0.2)
        ms = _timed(lambda: res.__setitem__("v", c.sss_verify(rows, scols, 1, 4)), steps, warm)
        f, sz = res["v"]
        sss[k] = {"verify_ms": ms, "similar": int(f.sum().item()), "lost_keys": c.failed,
                  "sizes_equal_reset": bool(torch.equal(sz, sizes0)) if c.failed == 0 else None}
        c.close()
    # ---- IR (the reference's ir_cuco comparison), a 1M-document index
    ioff, idocs = inputs.ir_index(num_docs=1_000_000, num_terms=100_000, cap=200_000, seed=SEED + 5)
    q, qoff = inputs.ir_queries(ioff, 20_000, seed=SEED + 6)
    qd, qod = torch.from_numpy(q).cuda(), torch.from_numpy(qoff).cuda()
    t = api.HashTables(ioff, idocs, 0.2, 5, 3)
    ikeys = t.probe_order(1)
    ir = {}
    ms = _timed(lambda: res.__setitem__("v", t.ir_query(qd, qod, 0)), steps, warm)
    total0 = int(res["v"][1].sum().item())
    ir["reset_normal"] = {"query_ms": ms, "results": total0}
    ms = _timed(lambda: res.__setitem__("v", t.ir_query(qd, qod, 1, term0_docs=ikeys)), steps, warm)
    ir["reset_hierarchical"] = {"query_ms": ms, "results": int(res["v"][1].sum().item())}
    t.close()
    # the same load factor with one-slot buckets (the IR kernels' probe width
    # is one lane per key): a probe's walk stays in one slot plane
    t = api.HashTables(ioff, idocs, 0.2, 1, 2)
    ikeys = t.probe_order(1)
    ms = _timed(lambda: res.__setitem__("v", t.ir_query(qd, qod, 1, term0_docs=ikeys)), steps, warm)
    ir["reset_hierarchical_b1"] = {"query_ms": ms, "results": int(res["v"][1].sum().item())}
    t.close()
    for k in kinds:
        c = api.CompetitorTables(ioff, idocs, k, 0.2)
        ms = _timed(lambda: res.__setitem__("v", c.ir_query(qd, qod)), steps, warm)
        ir[k] = {"query_ms": ms, "results": int(res["v"][1].sum().item()), "lost_keys": c.failed}
        c.close()
    tri = {v["triangles"] for v in tc.values() if v.get("lost_keys", 0) == 0}
    best = tc["reset_tiled_hierarchical"]
    return {
        "metric": "triangles/s (1 GPU), RESET tables beside the baseline directory's competitor tables",
        "value": best["triangles"] / (best["count_ms"] / 1000.0),
        "unit": "triangles/s",
        "n_gpus": 1, "steps": steps, "warmup": warm, "ms_per_step": best["count_ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic R-MAT (a=.57 b=.19 c=.19, seeded, labels scrambled)",
        "config": {"workload": f"TC on R-MAT scale {args.scale} (degree-oriented CSR, {E} edges); SSS on the "
                               f"{P} edges of a symmetric R-MAT scale {args.scale - 2} CSR at J >= 1/4; IR: 20k "
                               f"queries over a {int(ioff[-1])}-posting index (1M docs, 100k terms)",
                   "load_factor": 0.2, "bucket_size": 5, "parallelism": "single GPU"},
        "tc": tc, "sss": sss, "ir": ir,
        "parity": {"tc_counts_agree": len(tri) == 1,
                   "sss_counts_agree": len({v["similar"] for v in sss.values() if v.get("lost_keys", 0) == 0}) == 1,
                   "ir_counts_agree": len({v["results"] for v in ir.values() if v.get("lost_keys", 0) == 0}) == 1},
    }


def bench_pairs_shapes(args):
    """Launch shapes of the batched pair kernel (waves per CU, LDS table slots
    per wave, per-bin load factor and chunk, static or work-stealing schedule)
    on the headline 1M pairs: kernel
    ms (best of --steps launches, HIP events) per shape, all bit-checked.  A
    comparison line, not the headline metric."""
    import torch

    from reset_amd import api, inputs

    batch = inputs.make_pair_batch(args.pairs, seed=SEED)
    _sync_all()
    P = args.pairs
    n = int(batch["a_off"][-1]) + int(batch["b_off"][-1])
    # (waves/CU, table slots/wave, load factor, chunk, schedule: 0 static, k work stealing k pairs a take,
    #  unroll)
    shapes = [(16, 2048, 0.125, 255, 0, 2), (16, 2048, 0.125, 255, 1, 2), (16, 2048, 0.125, 255, 4, 2),
              (16, 2048, 0.125, 255, 8, 2), (16, 2048, 0.125, 255, 16, 2), (16, 2048, 0.125, 255, 32, 2),
              (16, 2048, 0.125, 255, 64, 2), (16, 2048, 0.125, 255, 8, 1), (16, 2048, 0.125, 255, 8, 4),
              (16, 2048, 0.25, 511, 0, 2), (16, 2048, 0.25, 511, 16, 2), (8, 4096, 0.125, 511, 0, 2),
              (8, 4096, 0.125, 511, 16, 2), (20, 1536, 0.25, 255, 0, 2), (20, 1536, 0.25, 255, 16, 2),
              (24, 1024, 0.25, 255, 0, 2), (24, 1024, 0.25, 255, 16, 2)]
    rows = []
    for wpc, slots, alpha, chunk, sched, unroll in shapes:
        opts = api.default_pairs_opts()
        opts.waves_per_cu, opts.lds_slots_per_wave, opts.schedule = wpc, slots, sched
        for i in range(opts.num_bins):
            opts.bins[i].load_factor, opts.bins[i].chunk, opts.bins[i].unroll = alpha, chunk, unroll
        plan = api.PairsPlan(P, opts)
        sizes = torch.empty(P, dtype=torch.int32, device="cuda")
        plan.run_async(batch["a_off"], batch["a"], batch["b_off"], batch["b"], sizes)
        plan.set_timing(True)
        ms = []
        for _ in range(max(1, args.steps)):
            plan.run_async(batch["a_off"], batch["a"], batch["b_off"], batch["b"], sizes)
            ms.append(plan.kernel_ms())
        plan.status()
        rows.append({"waves_per_cu": wpc, "lds_slots_per_wave": slots, "load_factor": alpha, "chunk": chunk,
                     "unroll": unroll, "schedule": f"steal {sched}" if sched else "static", "kernel_ms": min(ms), "alg_TBps": 4 * n / min(ms) / 1e9,
                     "bit_exact": bool(torch.equal(sizes, batch["k"]))})
        plan.close()
    best = min(rows, key=lambda r: r["kernel_ms"])
    return {"metric": "batched pair kernel ms per launch shape (1 GPU)", "value": best["kernel_ms"], "unit": "ms",
            "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": best["kernel_ms"],
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
            "data": "synthetic (the headline pairs)",
            "config": {"workload": f"batched pairwise intersection, {P} pairs, launch-shape comparison",
                       "pairs": P, "parallelism": "single GPU"},
            "shapes": rows}


HEADLINE_METRIC = "triangles/s & set-pair intersections/s (1 GPU); IR query p50/p99 at fixed QPS"
_HEADER = ("metric", "n_gpus", "steps", "warmup", "higher_is_better", "scaling", "vs_baseline", "dtype", "data")


def _progress(msg):
    """One line on stderr when a block of the default run ends."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def _block(d):
    return {k: v for k, v in d.items() if k not in _HEADER}


def bench_all(args):
    """The default line: the batched pairs (config 2, the headline value),
    then triangle counting at R-MAT scale 24 (config 3) and 16 (config 1)
    and the IR serving loop (config 5), each block with its own roofline and
    CPU baseline."""
    import torch

    t0 = time.perf_counter()
    out = bench_pairs(args)
    _progress(f"pairs block done ({time.perf_counter() - t0:.0f} s)")
    torch.cuda.empty_cache()
    tc24 = bench_tc(args, 24)
    _progress(f"tc scale-24 block done ({time.perf_counter() - t0:.0f} s)")
    torch.cuda.empty_cache()
    tc16 = bench_tc(args, 16)
    _progress(f"tc scale-16 block done ({time.perf_counter() - t0:.0f} s)")
    torch.cuda.empty_cache()
    ir = bench_ir(args)
    _progress(f"ir block done ({time.perf_counter() - t0:.0f} s)")
    out["metric"] = HEADLINE_METRIC
    out["headline"] = ("value: set-pair intersections/s of the 1M-pair batch; triangles/s of R-MAT scale 24 in "
                       "triangles_per_s (block tc); IR p50/p99 at fixed QPS in ir_latency_ms (block ir)")
    out["triangles_per_s"] = tc24["value"]
    out["ir_latency_ms"] = {"qps": args.qps, "p50": ir["latency_ms"]["p50"], "p99": ir["latency_ms"]["p99"]}
    out["tc"] = _block(tc24)
    out["tc_scale16"] = _block(tc16)
    out["ir"] = _block(ir)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="all",
                    choices=["all", "pairs", "tc", "sss", "ir", "ir_batch", "compare", "pairs_shapes"])
    ap.add_argument("--sets", type=int, default=1_000_000)
    ap.add_argument("--queries", type=int, default=20_000)
    ap.add_argument("--qps", type=float, default=2000.0, help="ir: the QPS whose latency heads the line")
    ap.add_argument("--qps-sweep", type=lambda x: [float(v) for v in x.split(",")],
                    default=[500.0, 1000.0, 2000.0, 5000.0], help="ir: arrival rates served")
    ap.add_argument("--serve-seconds", type=float, default=8.0,
                    help="ir: queries served per rate = rate x this (at most --queries, at least 2000)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="tc/ir: CPU baseline sample grown to this")
    ap.add_argument("--ir-budget-mb", type=int, default=4096)
    ap.add_argument("--ir-alpha", type=float, default=None,
                    help="ir: table load factor (default: the solver's pick for one lane per key)")
    ap.add_argument("--ir-bucket", type=int, default=None, help="ir: table bucket size (default: the solver's)")
    ap.add_argument("--ir-batch", type=int, default=256)
    ap.add_argument("--mixed-inflight", type=int, default=4,
                    help="IR streams of the mixed serving run (the lookups take one more)")
    ap.add_argument("--sss-plant", type=int, default=1000,
                    help="sss workload: near-duplicate sets planted at the end of the collection")
    ap.add_argument("--sss-lookups", type=int, default=0,
                    help="ir workload: also serve this many similarity lookups mixed with the queries")
    ap.add_argument("--sss-qps", type=float, default=500.0)
    ap.add_argument("--sss-sets", type=int, default=200_000)
    ap.add_argument("--pairs", type=int, default=1_000_000)
    ap.add_argument("--pair-lo", type=int, default=16, help="pairs: smallest set size")
    ap.add_argument("--pair-hi", type=int, default=65536, help="pairs: largest set size")
    ap.add_argument("--scale", type=int, default=None,
                    help="R-MAT scale: 24 for tc, 20 for compare (whose SSS part uses scale - 2)")
    ap.add_argument("--cpu-sample", type=int, default=50_000,
                    help="pairs (pairs workload) or sets (sss workload) in the CPU sample")
    ap.add_argument("--plain-sample", type=int, default=20_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="pairs: keep the default per-bin settings")
    ap.add_argument("--tune-sample", type=int, default=200_000, help="pairs: pairs the autotuner times on")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="also run a plain copy of this many MiB (counter calibration)")
    args = ap.parse_args(argv)
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on exactly one GPU (asked for --gpus {args.gpus}, WORLD_SIZE {world}); "
              "this build measures the single-GPU hot path only", file=sys.stderr)
        return 2
    import torch

    from reset_amd import lib

    if not torch.cuda.is_available():
        print("bench.py needs an MI355X GPU", file=sys.stderr)
        return 2
    lib.load()
    if args.copy_probe_mb:
        # a plain device copy of known size, profiled beside the measured kernel
        # to calibrate rocprofv3's FETCH_SIZE / WRITE_SIZE counters
        src = torch.ones(args.copy_probe_mb << 18, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.scale is None:
        args.scale = 20 if args.workload == "compare" else 24
    out = {"all": bench_all, "pairs": bench_pairs, "tc": bench_tc, "sss": bench_sss, "ir": bench_ir,
           "ir_batch": bench_ir_batch,
           "compare": bench_compare, "pairs_shapes": bench_pairs_shapes}[args.workload](args)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
