"""Serving form: resident star schemas and an open-loop Poisson client for
olap_server (continuous batching of star-join group-by queries).

Three resident datasets share the 4-dimension star of the plain benchmark
(date 2,556 / customer 30,000 / part 200,000 / supplier 2,000 rows) at 6M,
60M and 600M fact rows (about 0.12, 1.2 and 12 GB of fact columns). Each
arrival draws a dataset and one of a pool of random 1-4-dimension query
templates uniformly; inter-arrival times are exponential at the offered
rate and arrivals are submitted on schedule whatever the server's state
(open loop), so queueing delay shows in the latency.
"""
from __future__ import annotations

import threading
import time
from dataclasses import dataclass
from typing import List, Sequence

import numpy as np

from . import olapcore as oc
from . import workloads as wl

SERVE_ROWS = (6_000_000, 60_000_000, 600_000_000)


def serve_configs(rows: Sequence[int] = SERVE_ROWS) -> List[wl.StarConfig]:
    return [wl.scaled(wl.STAR_4DIM_600M, r, name=f"serve_4dim_{r // 1_000_000}M" if r >= 1_000_000
                      else f"serve_4dim_{r}") for r in rows]


def templates(n: int = 256, seed: int = 3, max_groups: int = 100_000):
    """n query templates [(steps, factor, G)] over the star's four tables."""
    return wl.batch_queries(wl.STAR_4DIM_600M, n, seed, max_groups=max_groups)


@dataclass
class Arrival:
    t: float          # scheduled time after start, seconds
    dataset: int      # index into the server's datasets
    template: int     # index into the template pool


def arrivals(rate: float, n: int, ndatasets: int, ntemplates: int, seed: int = 1) -> List[Arrival]:
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=n)
    t = np.cumsum(gaps) - gaps[0]
    ds = rng.integers(0, ndatasets, size=n)
    tp = rng.integers(0, ntemplates, size=n)
    return [Arrival(float(t[i]), int(ds[i]), int(tp[i])) for i in range(n)]


def run_open_loop(server: oc.Server, dataset_ids: Sequence[int], queries: Sequence[oc.BatchQuery],
                  plan: Sequence[Arrival], keep_results=False):
    """Submit `plan` on schedule from this thread while a collector thread
    waits for the results in submission order. Latency of a request = its
    submit lag behind schedule + the server's submit-to-result time.
    Returns a dict of q/s, p50/p99/max latency (ms) and, when asked, the
    results [(arrival index, sum, count)]: of every arrival when
    keep_results is True, of the arrivals i with keep_results(i) when it is
    a function."""
    keep = keep_results if callable(keep_results) else (lambda i: bool(keep_results))
    tickets: List[tuple] = []
    cv = threading.Condition()
    done = {"lat": [], "err": [], "results": [], "t_last": 0.0}
    n = len(plan)

    def collect():
        k = 0
        while k < n:
            with cv:
                while len(tickets) <= k:
                    cv.wait()
                i, tk, lag = tickets[k]
            try:
                s, c, lat_us = server.wait(tk, with_latency=True)
                done["lat"].append(lag + lat_us * 1e-6)
                if keep(i):
                    done["results"].append((i, s, c))
            except oc.OlapError as e:  # a failed batch: reported, not hidden
                done["err"].append((i, str(e)))
            done["t_last"] = time.perf_counter()
            k += 1

    th = threading.Thread(target=collect, daemon=True)
    th.start()
    t0 = time.perf_counter()
    for i, a in enumerate(plan):
        while True:
            now = time.perf_counter() - t0
            if now >= a.t:
                break
            if a.t - now > 2e-3:
                time.sleep(a.t - now - 1e-3)
        lag = time.perf_counter() - t0 - a.t
        tk = server.submit(dataset_ids[a.dataset], queries[a.template])
        with cv:
            tickets.append((i, tk, lag))
            cv.notify()
    th.join()
    span = max(done["t_last"] - t0, 1e-9)
    lat = np.sort(np.array(done["lat"])) * 1e3 if done["lat"] else np.zeros(1)
    out = {
        "offered_qps": n / max(plan[-1].t, 1e-9) if n > 1 else 0.0,
        "qps": len(done["lat"]) / span,
        "completed": len(done["lat"]),
        "failed": len(done["err"]),
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "max_ms": float(lat[-1]),
        "span_s": span,
    }
    if keep_results:
        out["results"] = done["results"]
    if done["err"]:
        out["errors"] = done["err"][:4]
    return out
