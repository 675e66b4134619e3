#!/usr/bin/env python
"""Benchmark of the MI355X star-join group-by path.

A step is one plain star-join group-by scan (the fused row-wise strategy by
default) over the whole device-resident fact table of the workload:
600M fact rows, 4 int32 FK columns + 1 int32 measure (12 GB), dimensions
date 2,556 / customer 30,000 / part 200,000 / supplier 2,000 rows with
selectivities 0.5 / 0.2 / 0.04 / 0.2 and 7 / 25 / 1 / 5 groups (G = 875).

Prints ONE JSON line (rank 0) with value = fact rows/s, the roofline of the
scan kernel (algorithmic bytes = rows x 20 B per launch over the HIP-event
average duration) and the CPU baseline timed on this host.
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
    ref = po.Reference()
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
    return {"value": n / t, "unit": "fact rows/s", "cores": cores, "kind": kind,
            "sample": f"first {n} fact rows of the workload, OLAPcore_vwm_dv, "
                      f"median of 5 runs after 1 warmup, {cores} pinned threads"}


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


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c3_4dim_600M")
    ap.add_argument("--strategy", default="rowwise")
    ap.add_argument("--cpu-rows", type=int, default=100_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the other strategies / configs")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="also run one device copy of MIB MiB (HBM-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the star-join path is measured on a single "
              "MI355X (see DESIGN.md for how it would shard); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    from memory_database_amd import olapcore as oc
    from memory_database_amd import workloads as wl

    cfg = wl.CONFIGS[args.config]
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
    log(f"autotune: shape {shape[0]} = {shape[1]} threads x {shape[2]} rows/thread "
        f"({shape[3]:.3f} ms)")

    def step():
        q.run(args.strategy, osum, ocnt, stream=s_handle)

    wall_ms, ev = time_steps(step, args.steps, args.warmup, stream)
    ms = statistics.mean(ev)
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

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(ds, args.cpu_rows)
    q.close()

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
                   "launch": {"threads": shape[1], "rows_per_thread": shape[2]},
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes": alg_bytes, "kernel_ms": ms},
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
