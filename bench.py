#!/usr/bin/env python
"""Benchmark of the MI355X star-join group-by path.

A step is one plain star-join group-by scan (the fused row-wise strategy by
default) over the whole device-resident fact table of the workload:
600M fact rows, 4 int32 FK columns + 1 int32 measure (12 GB), dimensions
date 2,556 / customer 30,000 / part 200,000 / supplier 2,000 rows with
selectivities 0.5 / 0.2 / 0.04 / 0.2 and 7 / 25 / 1 / 5 groups (G = 875).

Prints ONE JSON line (rank 0) with value = fact rows/s, the roofline of the
scan kernel (algorithmic bytes = rows x 20 B per launch over the HIP-event
average duration), the CPU baseline timed on this host, and beside them:
  batched  64 mixed queries (1-4 dimensions, 1-50% selectivity per
           dimension, 1..10^5 groups) over the same fact table in one fused
           scan replayed as a hipGraph, against the same 64 queries run as
           64 plain scans (queries/s of both forms and their ratio);
  served   an open-loop Poisson stream of random 1-4-dimension queries over
           resident datasets of 6M, 60M and 600M fact rows through the
           continuous-batching server (queries/s, p50/p99 latency per rate).
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "fact rows/s + % HBM peak, 4-dim star-join group-by; batched vs plain QPS, p99"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_steps(fn, steps, warmup, stream):
    """Warmup, then time exactly `steps` calls with HIP events on `stream`;
    returns (total_ms wall, per-step event ms list)."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a, b in evs:
        a.record(stream)
        fn()
        b.record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3
    return wall, [a.elapsed_time(b) for a, b in evs]


def cpu_baseline(ds, sample_rows, log_fn=log):
    """Time the CPU path on the first sample_rows fact rows of the workload
    (copied from the device): the reference's own OLAPcore_vwm_dv driver when
    oracle/_ref was built and its CPU pinning fits this process, else the
    oracle's restatement; all usable cores."""
    import numpy as np

    from oracle import pyoracle as po

    n = min(sample_rows, ds.cfg.rows)
    dims = ds.cfg.dims
    fks = [None] * 4
    dvs = [None] * 4
    for d in dims:
        fks[d.slot] = ds.fks[d.slot][:n].cpu().numpy()
        dvs[d.slot] = ds.dimvecs[d.slot].cpu().numpy()
    m1 = ds.m1[:n].cpu().numpy()
    m2 = np.zeros(n, dtype=np.int32)
    cores = po.available_cpus()
    plan = ds.plan
    orders = list(plan.orders)[: plan.dimvec_nums]
    factor = list(plan.factor)[: plan.dimvec_nums]
    kind = "port"
    ref = po.Reference(avx512=True)
    allowed = os.sched_getaffinity(0)
    use_ref = ref.ok and allowed == set(range(len(allowed)))
    if use_ref:
        kind = "reference"
        # the reference reads dimvec[fk] for every key in [0, D]; pad one NULL byte
        dvs_ref = []
        for t in range(4):
            if dvs[t] is None:
                dvs_ref.append(np.full(1, 127, dtype=np.int8))
            else:
                dvs_ref.append(np.concatenate([dvs[t], np.array([127], dtype=np.int8)]))
        fk_ref = [f if f is not None else np.zeros(1, dtype=np.int32) for f in fks]

        def run():
            return ref.olapcore_vwm_dv(dvs_ref, fk_ref, orders, factor, int(plan.group_nums), m1,
                                       m2, cores)
    else:
        star = po.Star(dvs, fks, m1, None)
        p = po.plan(orders, factor, int(plan.group_nums))

        def run():
            return po.run("vwm_dv", p, star, nthreads=cores)

    run()
    ts = []
    for _ in range(5):
        t0 = time.perf_counter()
        run()
        ts.append(time.perf_counter() - t0)
    t = statistics.median(ts)
    log_fn(f"cpu baseline ({kind}, {cores} threads): {n} rows in {t * 1e3:.1f} ms")
    build = ("reference sources built -O3 with the reference's AVX-512 flags (its makefile "
             "builds -O0)" if kind == "reference" and ref.avx512 else
             "reference sources built -O2 (this host lacks AVX-512; the reference builds -O0)"
             if kind == "reference" else "C restatement -O3")
    return {"value": n / t, "unit": "fact rows/s", "cores": cores, "kind": kind,
            "sample": f"first {n} fact rows of the workload, OLAPcore_vwm_dv ({build}), "
                      f"median of 5 runs after 1 warmup, {cores} pinned threads"}


def cpu_configs(ctx, stream, c4_rows=60_000_000):
    """The CPU path on the other configurations, on this host's cores: the
    1-dimension 6M-row star and the 4-dimension 60M-row star in full through
    the reference's own OLAPcore_vwm_dv and OLAPcore_rowwise drivers, and the
    64 batched queries run one after the other by the C restatement over the
    first c4_rows rows of the 600M-row star (their group codes exceed the
    reference's int8 vectors)."""
    import numpy as np
    import torch

    from memory_database_amd import workloads as wl
    from oracle import pyoracle as po

    cores = po.available_cpus()
    ref = po.Reference(avx512=True)
    use_ref = ref.ok and os.sched_getaffinity(0) == set(range(len(os.sched_getaffinity(0))))
    out = {}
    for cfg in (wl.STAR_1DIM_6M, wl.STAR_4DIM_60M):
        ds = wl.DeviceStar(ctx, cfg, stream=stream.cuda_stream)
        torch.cuda.synchronize()
        dvs, fks = [None] * 4, [None] * 4
        for d in cfg.dims:
            dvs[d.slot] = ds.dimvecs[d.slot].cpu().numpy()
            fks[d.slot] = ds.fks[d.slot].cpu().numpy()
        m1 = ds.m1.cpu().numpy()
        m2 = np.zeros_like(m1)
        plan = ds.plan
        orders = list(plan.orders)[: plan.dimvec_nums]
        factor = list(plan.factor)[: plan.dimvec_nums]
        del ds
        row = {"rows": cfg.rows, "cores": cores}
        for strat in ("vwm_dv", "rowwise"):
            if use_ref:
                dvr = [np.concatenate([d, np.array([127], np.int8)]) if d is not None
                       else np.full(1, 127, np.int8) for d in dvs]
                fkr = [f if f is not None else np.zeros(1, np.int32) for f in fks]
                fn = getattr(ref, "olapcore_" + strat)

                def run():
                    return fn(dvr, fkr, orders, factor, int(plan.group_nums), m1, m2, cores)
            else:
                star = po.Star(dvs, fks, m1, None)
                p = po.plan(orders, factor, int(plan.group_nums))

                def run():
                    return po.run(strat, p, star, nthreads=cores)
            run()
            ts = []
            for _ in range(3):
                t0 = time.perf_counter()
                run()
                ts.append(time.perf_counter() - t0)
            t = statistics.median(ts)
            row[strat] = {"ms": t * 1e3, "rows_per_s": cfg.rows / t}
        row["kind"] = "reference" if use_ref else "port"
        row["build"] = ("-O3 + the reference's AVX-512 flags" if use_ref and ref.avx512 else "-O2" if use_ref
                        else "C restatement -O3")
        log(f"cpu {cfg.name}: {row}")
        out[cfg.name] = row
    # C4: 64 queries in sequence (the plain scans' CPU equivalent)
    cfg = wl.scaled(wl.STAR_4DIM_600M, c4_rows, name=f"c4_sample_{c4_rows}")
    bs = wl.DeviceBatchStar(ctx, cfg)
    torch.cuda.synchronize()
    filt = [None if f is None else f.cpu().numpy() for f in bs.filters]
    attr = [None if a is None else a.cpu().numpy() for a in bs.attrs]
    fks = [None if f is None else f.cpu().numpy() for f in bs.fks]
    m1 = bs.m1.cpu().numpy()
    del bs
    qs = wl.batch_queries(wl.STAR_4DIM_600M, 64, seed=7)
    t_all = 0.0
    for steps, factor, G in qs:
        dvs = [None] * 4
        for (t, lo, hi, g) in steps:
            dvs[t] = po.build_dimvec(filt[t], attr[t], lo, hi, 1 if g > 0 else 0, 4, group_mod=g)
        p = po.plan([s[0] for s in steps], factor, G)
        star = po.Star(dvs, fks, m1, None)
        t0 = time.perf_counter()
        po.run("rowwise", p, star, nthreads=cores)
        t_all += time.perf_counter() - t0
    out["batch_64q_sequential"] = {"rows": c4_rows, "queries": 64, "ms": t_all * 1e3,
                                   "qps": 64 / t_all, "cores": cores, "kind": "port",
                                   "note": "C restatement (row-wise, int32 vectors), one query after "
                                           f"another over the first {c4_rows} rows of the 600M-row star"}
    log(f"cpu 64 queries over {c4_rows} rows: {t_all * 1e3:.0f} ms")
    return out


def source_digest() -> str:
    """First 16 hex digits of the SHA-256 of the HIP kernel sources; a
    profile summary records the digest of the sources it was taken on."""
    import hashlib

    h = hashlib.sha256()
    csrc = os.path.join(ROOT, "memory_database_amd", "csrc")
    for f in sorted(os.listdir(csrc)):
        if f.endswith((".hip", ".h")):
            h.update(open(os.path.join(csrc, f), "rb").read())
    return h.hexdigest()[:16]


def profiled_traffic(kernel: str, workload: str, key: str = "HBM traffic"):
    """HBM bytes per dispatch of `kernel` on `workload` from the committed
    rocprofv3 summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated), the
    summary's path, and whether it was taken on the current kernel sources;
    (None, None, None) without a summary."""
    import re

    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}.md")
    try:
        text = open(path).read()
    except OSError:
        return None, None, None
    m = re.search(re.escape(key) + r" ([0-9.]+) GB/dispatch", text)
    if not m:
        return None, None, None
    d = re.search(r"kernel sources digest ([0-9a-f]{16})", text)
    return float(m.group(1)) * 1e9, os.path.relpath(path, ROOT), bool(d and d.group(1) == source_digest())


def roofline(alg_bytes: float, ms: float, traffic, source, current):
    """Roofline of one launch: achieved = HBM bytes moved (the profile's
    counters) / the HIP-event kernel time, as a fraction of the HBM peak;
    the algorithmic bytes over the same time are reported beside it."""
    moved = traffic if traffic else None
    gbps = (moved if moved else alg_bytes) / (ms / 1e3) / 1e9
    return {"bound": "hbm", "achieved": gbps, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": gbps / HBM_PEAK_GBS, "traffic": moved,
            "basis": "HBM bytes moved per launch (FETCH_SIZE + WRITE_SIZE, copy-calibrated) / "
                     "HIP-event kernel time" if moved else
                     "algorithmic bytes / HIP-event kernel time (no counter profile)",
            "traffic_source": source, "traffic_on_current_sources": current,
            "algorithmic_bytes": alg_bytes, "kernel_ms": ms,
            "effective_GBps": alg_bytes / (ms / 1e3) / 1e9,
            "effective_frac": alg_bytes / (ms / 1e3) / 1e9 / HBM_PEAK_GBS}


def bench_sweep(ctx, stream, rows, strategy="rowwise"):
    """The plain scan across the sweep of the 4-dimension star (Zipf skew of
    the foreign keys x total groups x per-dimension selectivity), each
    configuration autotuned, then timed as the headline scan is."""
    import torch

    from memory_database_amd import workloads as wl

    sh = stream.cuda_stream
    out = []
    fact = None
    theta = None
    for cfg in wl.sweep_configs(rows):
        if cfg.zipf_theta != theta:
            fact, theta = None, cfg.zipf_theta
        with torch.cuda.stream(stream):
            ds = wl.DeviceStar(ctx, cfg, stream=sh, fact=fact)
        torch.cuda.synchronize()
        fact = fact or ds
        G = int(ds.plan.group_nums)
        q = ctx.query(ds.plan, ds.cols)
        osum = torch.empty(G, dtype=torch.int64, device="cuda")
        ocnt = torch.empty(G, dtype=torch.int64, device="cuda")
        q.autotune(strategy, osum, ocnt, stream=sh)
        _, ev = time_steps(lambda: q.run(strategy, osum, ocnt, stream=sh), 10, 2, stream)
        ms = statistics.mean(ev)
        alg = ds.fact_bytes()
        th, rp, agg = q.launch_info()
        row = {"theta": cfg.zipf_theta, "groups": G, "sele_pct": cfg.dims[0].pred_hi,
               "ms": round(ms, 4), "rows_per_s": rows / (ms / 1e3),
               "effective_GBps": round(alg / (ms / 1e3) / 1e9, 1),
               "passing_rows": int(ocnt.sum().item()),
               "launch": {"threads": th, "rows_per_thread": rp, "aggregation": agg, **q.pipeline()}}
        log(f"sweep {cfg.name}: {row}")
        out.append(row)
        q.close()
        del ds, osum, ocnt
    return {"rows": rows, "strategy": strategy, "note": "every dimension at the same selectivity; "
            "G split over date/customer/part/supplier; foreign keys Zipf(theta) over each "
            "dimension's rows", "configs": out}


def bench_batched(ctx, ds, stream, nq=64, seed=7):
    """The batched form on the plain workload's fact table: nq seeded mixed
    queries in one fused scan (autotuned, then captured and replayed as a
    hipGraph) against the same queries as nq plain scans."""
    import torch

    from memory_database_amd import olapcore as oc
    from memory_database_amd import workloads as wl

    sh = stream.cuda_stream
    bs = wl.DeviceBatchStar(ctx, ds.cfg, fact=ds, stream=sh)
    torch.cuda.synchronize()
    qs = wl.batch_queries(ds.cfg, nq, seed)
    b = ctx.batch(bs.star, [oc.BatchQuery.make(*q) for q in qs])
    osum = torch.empty(b.total_slots, dtype=torch.int64, device="cuda")
    ocnt = torch.empty_like(osum)
    cfg, threads, rpt, kib, staged, tune_ms = b.autotune(osum, ocnt, sh)
    b.capture(osum, ocnt, sh)
    _, ev = time_steps(lambda: b.launch(sh), 5, 1, stream)
    batch_ms = statistics.mean(ev)
    # every query's slots of the last timed replay must equal its plain scan
    bsum, bcnt = osum.clone(), ocnt.clone()
    plain_ms = 0.0
    for i, (steps, factor, G) in enumerate(qs):
        plan, cols = bs.plain_query(steps, factor, G, stream=sh)
        q = ctx.query(plan, cols)
        ps = torch.empty(G, dtype=torch.int64, device="cuda")
        pc = torch.empty_like(ps)
        _, e = time_steps(lambda: q.run("rowwise", ps, pc, stream=sh), 2, 1, stream)
        plain_ms += statistics.mean(e)
        q.close()
        lo, hi = int(b.offsets[i]), int(b.offsets[i + 1])
        if not (torch.equal(bsum[lo:hi], ps) and torch.equal(bcnt[lo:hi], pc)):
            raise RuntimeError(f"batched query {i} differs from its plain scan")
    tables = set(t for steps, _, _ in qs for (t, _, _, _) in steps)
    rows = ds.cfg.rows
    union_bytes = rows * 4 * (len(tables) + 1)
    per_query_bytes = sum(rows * 4 * (len(steps) + 1) for steps, _, _ in qs)
    events = b.expected_events()
    b.close()
    log(f"batched: {nq} queries {batch_ms:.2f} ms (config {cfg}: {threads} threads, {rpt} rows/thread, "
        f"{kib} KiB LDS tables, staged={staged}); plain {plain_ms:.2f} ms; x{plain_ms / batch_ms:.2f}; "
        f"all {nq} queries equal their plain scans")
    traffic, src, current = profiled_traffic(oc.config_kernel(cfg), f"{nq}q_{ds.cfg.name}",
                                             key="batch HBM traffic")
    return {
        "verified": f"every slot of the {nq} queries equals its plain row-wise scan",
        "roofline": roofline(union_bytes, batch_ms, traffic, src, current),
        "queries": nq, "seed": seed, "batch_ms": batch_ms, "plain_ms_total": plain_ms,
        "qps_batched": nq / (batch_ms / 1e3), "qps_plain": nq / (plain_ms / 1e3),
        "speedup": plain_ms / batch_ms,
        "matches_per_row": events,
        "config": {"id": cfg, "kernel": oc.config_kernel(cfg), "threads": threads, "rows_per_thread": rpt,
                   "lds_table_kib": kib, "aggregation": ["atomics", "two-pass staged", "one-pass staged"][staged]},
        "union_bytes": union_bytes, "union_GBps": union_bytes / (batch_ms / 1e3) / 1e9,
        "effective_GBps": per_query_bytes / (batch_ms / 1e3) / 1e9,
        "timing": "hipGraph replay, HIP events, mean of 5 after 1 warmup; plain = sum of the "
                  "64 plain row-wise scans, each mean of 2 after 1 warmup",
    }


def bench_served(ctx, ds, rates, seconds, streams=2):
    """Open-loop Poisson arrivals through olap_server over resident datasets
    of 6M, 60M and 600M fact rows (the 600M one is the plain workload's)."""
    import torch

    from memory_database_amd import olapcore as oc
    from memory_database_amd import serving
    from memory_database_amd import workloads as wl

    stars = []
    for c in serving.serve_configs():
        if c.rows == ds.cfg.rows:
            stars.append(wl.DeviceBatchStar(ctx, c, fact=ds))
        else:
            stars.append(wl.DeviceBatchStar(ctx, c))
    torch.cuda.synchronize()
    qs = [oc.BatchQuery.make(*t) for t in serving.templates(256, seed=3)]
    srv = ctx.server(max_batch=64, streams=streams, memory_budget=32 << 30)
    ids = [srv.add_dataset(s.star) for s in stars]
    # warm-up: full batches on every dataset for every worker, so the
    # workers' grow-only staging has reached its size before any timing
    srv.pause(True)
    warm = [srv.submit(d, qs[i % len(qs)]) for d in ids for i in range(64 * streams)]
    srv.pause(False)
    for t in warm:
        srv.wait(t)
    serving.run_open_loop(srv, ids, qs, serving.arrivals(200.0, 60, len(ids), len(qs), seed=99))
    out = []
    samples = []  # (dataset, template, sum, count) of every 97th arrival (or sparser)
    for r in rates:
        n = int(max(100, min(r * seconds, 30000)))
        srv.reset_stats()
        plan = serving.arrivals(r, n, len(ids), len(qs), seed=1)
        every = max(97, n // 64)  # at most ~64 sampled results a rate
        res = serving.run_open_loop(srv, ids, qs, plan, keep_results=lambda i: i % every == 0)
        for i, su, co in res.pop("results"):
            samples.append((plan[i].dataset, plan[i].template, su, co))
        st = srv.stats()
        row = {"offered_qps": round(res["offered_qps"], 1), "requests": n, "qps": round(res["qps"], 1),
               "p50_ms": round(res["p50_ms"], 3), "p99_ms": round(res["p99_ms"], 3),
               "max_ms": round(res["max_ms"], 3), "mean_batch": round(st["mean_batch"], 2),
               "window_s": round(plan[-1].t, 3), "span_s": round(res["span_s"], 3), "failed": res["failed"]}
        log(f"served at {r:g} q/s: {row}")
        out.append(row)
    srv.close()
    # the sampled results against the same requests served alone
    import numpy as np

    alone = ctx.server(max_batch=1, streams=1, memory_budget=32 << 30)
    aids = [alone.add_dataset(s.star) for s in stars]
    per_ds = [0] * len(stars)
    for d, t, su, co in samples:
        s1, c1 = alone.wait(alone.submit(aids[d], qs[t]))
        if not (np.array_equal(su, s1) and np.array_equal(co, c1)):
            raise RuntimeError(f"served result of template {t} on dataset {d} differs from alone")
        per_ds[d] += 1
    alone.close()
    log(f"served: {len(samples)} sampled results equal the same requests served alone {per_ds}")
    return {"datasets_rows": [s.cfg.rows for s in stars], "templates": len(qs), "workers": streams,
            "max_batch": 64, "rates": out,
            "verified": f"{len(samples)} sampled results (every 97th arrival or sparser, per dataset {per_ds}) "
                        "equal the same requests served alone",
            "latency": "scheduled arrival to result on the host (submit lag + server time)"}


def bench_q5(ctx, stream, sf=100.0, region="ASIA", date=19940101):
    """TPC-H Q5 over a TPC-H-shaped snowflake at scale factor sf (6M x sf
    lineitems): the dimension vectors built with general predicates, then
    the fused probe (Tri_table_join), timed apart."""
    import torch

    from memory_database_amd import workloads as wl

    sh = stream.cuda_stream
    cfg = wl.Q5Config(sf)
    with torch.cuda.stream(stream):
        dq = wl.DeviceQ5(ctx, cfg, stream=sh)
    torch.cuda.synchronize()
    out_s = torch.empty(25, dtype=torch.int64, device="cuda")
    out_c = torch.empty(25, dtype=torch.int64, device="cuda")
    vecs = {}

    def build():
        vecs["v"] = dq.vectors(ctx, region, date, code_bytes=1, stream=sh)

    _, evb = time_steps(build, 5, 1, stream)
    sv, ov = vecs["v"]
    _, evp = time_steps(lambda: dq.run(ctx, sv, ov, out_s, out_c, stream=sh), 10, 2, stream)
    ms = statistics.mean(evp)
    alg = dq.fact_bytes()
    rows = cfg.lineitems
    nat = {n: int(c) for n, c in zip(range(25), out_c.cpu().tolist()) if c}
    log(f"Q5 sf={sf:g}: vectors {statistics.mean(evb):.3f} ms, probe {ms:.3f} ms over {rows} lineitems "
        f"({rows / ms / 1e6:.1f} G rows/s), {sum(nat.values())} matching rows in {len(nat)} nations")
    return {"sf": sf, "lineitems": rows, "region": region, "date": date,
            "vectors_ms": statistics.mean(evb), "probe_ms": ms, "rows_per_s": rows / (ms / 1e3),
            "effective_GBps": alg / (ms / 1e3) / 1e9, "algorithmic_bytes": alg,
            "matching_rows": sum(nat.values()), "nations": len(nat),
            "revenue_units": "cents x percent (10^4 x the reference's double revenue)",
            "timing": "HIP events, mean of 10 probes after 2 warmups; vectors mean of 5"}


def bench_substeps(ctx, stream):
    """The sub-step operators at the reference's own sizes: vector_group over
    600M rows for 2^5..2^26 groups (code/group, log/group/groupalgo_test.tsv
    times beside), the dense-vector join probe of 2^30 keys into 2^10..2^26
    entries (code/join)."""
    import torch

    sh = stream.cuda_stream
    rows = 600_000_000
    key = torch.empty(rows, dtype=torch.int32, device="cuda")
    m1 = torch.empty(rows, dtype=torch.int32, device="cuda")
    m2 = torch.empty(rows, dtype=torch.int32, device="cuda")
    ctx.gen_uniform(m1, 3, 1, 0, 1000, stream=sh)
    ctx.gen_uniform(m2, 3, 2, 0, 1000, stream=sh)
    # the reference's published single-thread times (ms, -O0), log/group/groupalgo_test.tsv
    published = {5: 7077, 8: 2838, 10: 2746, 12: 2414, 16: 2457, 18: 3423, 20: 4292, 22: 5014,
                 24: 11102, 26: 14438}
    group = []
    for k in sorted(published):
        G = 1 << k
        ctx.gen_uniform(key, 3, 10 + k, 1, G + 1, stream=sh)
        out = torch.empty(G, dtype=torch.int64, device="cuda")
        _, ev = time_steps(lambda: ctx.vector_group(key, m1, m2, out, stream=sh), 3, 1, stream)
        ms = statistics.mean(ev)
        group.append({"log2_groups": k, "ms": round(ms, 3), "rows_per_s": rows / (ms / 1e3),
                      "reference_cpu_ms": published[k]})
        log(f"vector_group 600M rows, 2^{k} groups: {ms:.3f} ms (reference CPU {published[k]} ms)")
    del key, m1, m2
    srows = 1 << 30
    skey = torch.empty(srows, dtype=torch.int32, device="cuda")
    join = []
    for k in (10, 15, 20, 26):
        maxid = 1 << k
        ctx.gen_uniform(skey, 4, k, 1, maxid + 1, stream=sh)
        rkey = torch.randperm(maxid, device="cuda", dtype=torch.int64).to(torch.int32) + 1
        vec = torch.zeros(maxid + 1, dtype=torch.int32, device="cuda")
        ctx.join_build(rkey, torch.ones_like(rkey), vec, stream=sh)
        res = {}

        def probe():
            res["t"] = ctx.join_probe(skey, vec, stream=sh)

        _, ev = time_steps(probe, 3, 1, stream)
        ms = statistics.mean(ev)
        assert res["t"] == (srows, 0)
        join.append({"log2_build_rows": k, "probe_rows": srows, "ms": round(ms, 3),
                     "rows_per_s": srows / (ms / 1e3)})
        log(f"join probe 2^30 keys into 2^{k} entries: {ms:.3f} ms")
    return {"vector_group": group, "join_probe": join,
            "timing": "HIP events around each synchronising call, mean of 3 after 1 warmup"}


def host_info():
    info = {}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core",
                             "NUMA node(s)", "CPU(s)"):
                info[k.strip()] = v.strip()
    except Exception:
        pass
    return info


def bench_sf100_queries(repeat: int, only=None):
    """The reference's ten recorded GPU OLAPcore runs at SF=100 (600M rows,
    its own gen_data with glibc rand() seed 1, the byte past each dimension
    vector 0), through the drop-in driver: every strategy's best of `repeat`
    runs after a warm one (HIP events), its count checked against the
    reference's recorded total, beside the time the reference printed for
    the same run and strategy. Host data generation takes ~20 s a query."""
    import subprocess
    import tempfile

    here = os.path.dirname(os.path.abspath(__file__))
    fixture = json.load(open(os.path.join(here, "tests", "golden", "known_totals_sf100.json")))
    driver = os.path.join(here, "memory_database_amd", "olapcore_test")
    out = []
    for case in fixture["cases"]:
        if only and case["sqlnum"] not in only:
            continue
        with tempfile.TemporaryDirectory() as tmp:
            js = os.path.join(tmp, "out.jsonl")
            cmd = [driver] + [f"--{k}={v}" for k, v in case["flags"].items()] + [
                f"--sqlnum={case['sqlnum']}", "--gen=reference", "--pad=0", "--strategy=all",
                f"--repeat={repeat}", f"--json={js}"]
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
            if r.returncode != 0:
                raise RuntimeError(f"driver failed on query {case['sqlnum']}: {r.stderr[-400:]}")
            recs = [json.loads(line) for line in open(js)]
        ours = {x["strategy"]: x["ms"] for x in recs}
        counts_ok = all(x["count"] == case["count"] for x in recs)
        pub = case["published_ms"]
        row = {"query": f"Q{case['sqlnum'] // 10}.{case['sqlnum'] % 10}", "count": case["count"],
               "counts_match_recorded": counts_ok, "ms": ours, "published_ms": pub,
               "best_ms": min(ours.values()), "published_best_ms": min(pub.values()),
               "best_speedup": min(pub.values()) / min(ours.values())}
        log(f"{row['query']}: best {row['best_ms']:.3f} ms ({min(ours, key=ours.get)}) against the "
            f"published {row['published_best_ms']:.3f} ms ({min(pub, key=pub.get)}): "
            f"x{row['best_speedup']:.1f}; counts {'match' if counts_ok else 'DIFFER'}; {ours}")
        out.append(row)
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="star_4dim_600M")
    ap.add_argument("--strategy", default="rowwise")
    ap.add_argument("--cpu-rows", type=int, default=100_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the other strategies / configs")
    ap.add_argument("--no-sweep", action="store_true", help="skip the skew/groups/selectivity sweep")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched form")
    ap.add_argument("--no-serve", action="store_true", help="skip the served form")
    ap.add_argument("--no-q5", action="store_true", help="skip TPC-H Q5 at scale factor 100")
    ap.add_argument("--no-substeps", action="store_true", help="skip the sub-step operators")
    ap.add_argument("--serve-rates", default="50,500,1000,2000,3000,5000")
    ap.add_argument("--serve-seconds", type=float, default=6.0,
                    help="arrival window per offered rate (q/s counts the drain after the last arrival)")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="also run one device copy of MIB MiB (HBM-counter calibration)")
    ap.add_argument("--sf100-queries", action="store_true",
                    help="only run the reference's ten recorded SF=100 queries through the drop-in "
                         "driver (minutes of host data generation) and print one JSON line")
    ap.add_argument("--sf100-repeat", type=int, default=5)
    args = ap.parse_args()
    if args.sf100_queries:
        rows = bench_sf100_queries(args.sf100_repeat)
        import math
        gm = math.exp(sum(math.log(r["best_speedup"]) for r in rows) / max(len(rows), 1))
        print(json.dumps({"metric": "SF=100 recorded queries: best strategy ms vs the reference's printed ms",
                          "queries": rows, "geomean_best_speedup": gm,
                          "all_counts_match": all(r["counts_match_recorded"] for r in rows),
                          "timing": f"best of {args.sf100_repeat} after a warm run, HIP events, per strategy",
                          "data": "reference gen_data (glibc rand() seed 1), 600M rows"}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the star-join path is measured on a single "
              "MI355X (multi-GPU sharding is a design, not a measured path here); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from memory_database_amd import olapcore as oc
    from memory_database_amd import workloads as wl

    cfg = wl.config_by_name(args.config)
    ctx = oc.Context(0)
    info = ctx.info()
    stream = torch.cuda.Stream()
    s_handle = stream.cuda_stream
    log(f"device {info}; generating {cfg.name} ({cfg.rows} rows)")
    with torch.cuda.stream(stream):
        ds = wl.DeviceStar(ctx, cfg, stream=s_handle)
    torch.cuda.synchronize()
    G = int(ds.plan.group_nums)
    q = ctx.query(ds.plan, ds.cols)
    osum = torch.empty(G, dtype=torch.int64, device="cuda")
    ocnt = torch.empty(G, dtype=torch.int64, device="cuda")

    # the autotuner times every launch shape once (untimed) and keeps the best
    shape = q.autotune(args.strategy, osum, ocnt, stream=s_handle)
    th, rp, agg = q.launch_info()
    log(f"autotune: shape {shape[0]} = {th} threads x {rp} rows/thread, {agg} "
        f"({shape[3]:.3f} ms)")

    def step():
        q.run(args.strategy, osum, ocnt, stream=s_handle)

    wall_ms, ev = time_steps(step, args.steps, args.warmup, stream)
    ms = statistics.mean(ev)
    th, rp, agg = q.launch_info()
    launch = {"threads": th, "rows_per_thread": rp, "aggregation": agg, **q.pipeline()}
    rows_per_s = cfg.rows / (ms / 1e3)
    alg_bytes = ds.fact_bytes()
    achieved = alg_bytes / (ms / 1e3) / 1e9
    passing = int(ocnt.sum().item())
    log(f"{cfg.name} {args.strategy}: {ms:.3f} ms/scan (wall {wall_ms / args.steps:.3f}), "
        f"{rows_per_s / 1e9:.1f} G rows/s, {achieved:.0f} GB/s algorithmic, {passing} passing rows")

    extra = {}
    if not args.no_extra:
        per = {}
        for s in oc.STRATEGIES:
            _, e = time_steps(lambda: q.run(s, osum, ocnt, stream=s_handle), 5, 1, stream)
            per[s] = round(statistics.mean(e), 4)
            log(f"  strategy {s}: {per[s]:.3f} ms")
        extra["strategies_ms"] = per

    if args.copy_calib:
        # a plain copy of known size: its FETCH_SIZE / WRITE_SIZE per dispatch
        # calibrate the counters read for the scan kernel
        nbytes = args.copy_calib << 20
        src = torch.empty(nbytes // 4, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        src.fill_(1)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: {nbytes} bytes read + written")
        del src, dst

    if not args.no_batch:
        extra["batched"] = bench_batched(ctx, ds, stream)
    if not args.no_sweep:
        extra["sweep"] = bench_sweep(ctx, stream, cfg.rows, args.strategy)
    if not args.no_serve:
        rates = [float(r) for r in args.serve_rates.split(",") if r]
        extra["served"] = bench_served(ctx, ds, rates, args.serve_seconds)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(ds, args.cpu_rows)
    q.close()
    if not args.no_cpu and not args.no_extra:
        del ds
        torch.cuda.empty_cache()
        ds = None
        cpu["configs"] = cpu_configs(ctx, stream)
    if not args.no_q5 or not args.no_substeps:
        # the sub-steps and Q5 need room: the star's fact columns are done with
        ds = None
        torch.cuda.empty_cache()
    if not args.no_q5:
        extra["tpch_q5"] = bench_q5(ctx, stream)
    if not args.no_substeps:
        extra["substeps"] = bench_substeps(ctx, stream)

    # the row-wise kernel the autotuned shape launched names the profile
    kernel = None
    if args.strategy == "rowwise":
        kernel = ("k_rowwise_queue" if launch.get("survivor_queue") else
                  "k_rowwise_pipe" if launch.get("fetch_ahead") else "k_rowwise")
    traffic, traffic_src, current = (profiled_traffic(kernel, cfg.name) if kernel
                                     else (None, None, None))

    rec = {
        "metric": METRIC,
        "value": rows_per_s,
        "unit": "fact rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic (counter-based SplitMix64 columns generated on the device)",
        "config": {"workload": cfg.name, "fact_rows": cfg.rows,
                   "dims": [{"slot": d.slot, "rows": d.rows, "sele": d.sele, "groups": d.groups}
                            for d in cfg.dims],
                   "measure": [cfg.m_lo, cfg.m_hi], "group_nums": G, "strategy": args.strategy,
                   "launch": launch,
                   "parallelism": "single GPU"},
        "roofline": dict(roofline(alg_bytes, ms, traffic, traffic_src, current),
                         note="algorithmic bytes count every FK and measure byte; the scan's early "
                              "exit loads later columns only for cache lines holding a live row, "
                              "so the bytes moved are below them"),
        "cpu_baseline": cpu,
        "passing_rows": passing,
        "wall_ms_per_step": wall_ms / args.steps,
        "device": info,
        "host": host_info(),
    }
    rec.update(extra)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
