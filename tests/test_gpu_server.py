"""The serving loop on the GPU: requests served alone and in a mix of
concurrent fused batches give identical results, equal to the oracle; an
open-loop Poisson run over resident datasets completes every request."""
import numpy as np
import pytest

from memory_database_amd import olapcore as oc
from memory_database_amd import serving
from memory_database_amd import workloads as wl
from tests.test_gpu_batch import host_star, oracle_query

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def test_server_alone_equals_mix_equals_oracle(ctx, ncpu):
    cfgs = [wl.scaled(wl.STAR_4DIM_600M, r, name=f"srv_{r}") for r in (1_000_003, 300_007)]
    dstars = [wl.DeviceBatchStar(ctx, c) for c in cfgs]
    hosts = [host_star(c, c.rows) for c in cfgs]
    tmpl = wl.batch_queries(wl.STAR_4DIM_600M, 48, seed=21, max_groups=20_000)
    rng = np.random.default_rng(4)
    ds = [int(x) for x in rng.integers(0, 2, len(tmpl))]
    qs = [oc.BatchQuery.make(*t) for t in tmpl]

    _say("alone: max_batch 1, one request at a time")
    alone = ctx.server(max_batch=1, streams=1, memory_budget=1 << 30)
    ids = [alone.add_dataset(d.star) for d in dstars]
    got_alone = [alone.wait(alone.submit(ids[d], q)) for d, q in zip(ds, qs)]
    alone.close()

    _say("mix: 48 requests admitted together over 2 workers")
    mix = ctx.server(max_batch=64, streams=2, memory_budget=1 << 30)
    ids = [mix.add_dataset(d.star) for d in dstars]
    mix.pause(True)
    tickets = [mix.submit(ids[d], q) for d, q in zip(ds, qs)]
    mix.pause(False)
    got_mix = [mix.wait(t) for t in tickets]
    st = mix.stats()
    _say(f"  stats {st}")
    assert st["completed"] == len(qs) and st["failed"] == 0 and st["batches"] <= 4
    mix.close()

    for i, (d, t) in enumerate(zip(ds, tmpl)):
        np.testing.assert_array_equal(got_mix[i][0], got_alone[i][0], err_msg=f"request {i} sums")
        np.testing.assert_array_equal(got_mix[i][1], got_alone[i][1], err_msg=f"request {i} counts")
        if i % 4 == 0:  # a quarter against the oracle (each costs a CPU star join)
            es, ec = oracle_query(hosts[d], *t, ncpu)
            np.testing.assert_array_equal(got_mix[i][0], es, err_msg=f"request {i} vs oracle")
            np.testing.assert_array_equal(got_mix[i][1], ec, err_msg=f"request {i} vs oracle")


def test_server_staged_batch_equals_plain_batch(ctx):
    """A batch heavy enough for the staged aggregation gives the same result
    per request as olap_batch run directly with the atomics layout."""
    import torch

    cfg = wl.scaled(wl.STAR_4DIM_600M, 40_000_000, name="srv_staged")
    dstar = wl.DeviceBatchStar(ctx, cfg)
    tmpl = wl.batch_queries(wl.STAR_4DIM_600M, 64, seed=33)
    qs = [oc.BatchQuery.make(*t) for t in tmpl]
    b = ctx.batch(dstar.star, qs)
    _say(f"expected matches {b.expected_events() * cfg.rows:.3g}")
    osum = torch.empty(b.total_slots, dtype=torch.int64, device="cuda")
    ocnt = torch.empty(b.total_slots, dtype=torch.int64, device="cuda")
    b.set_config(0)
    b.run(osum, ocnt)
    torch.cuda.synchronize()
    es, ec = osum.cpu().numpy(), ocnt.cpu().numpy()
    srv = ctx.server(max_batch=64, streams=2, memory_budget=4 << 30)
    d = srv.add_dataset(dstar.star)
    srv.pause(True)
    tickets = [srv.submit(d, q) for q in qs]
    srv.pause(False)
    for i, t in enumerate(tickets):
        s, c = srv.wait(t)
        lo, hi = b.offsets[i], b.offsets[i + 1]
        np.testing.assert_array_equal(s, es[lo:hi], err_msg=f"request {i} sums")
        np.testing.assert_array_equal(c, ec[lo:hi], err_msg=f"request {i} counts")
    srv.close()
    b.close()


def test_open_loop_poisson_small(ctx):
    """An open-loop Poisson run at 500 q/s over two resident datasets: every
    request completes; a sample equals the same request served alone."""
    cfgs = serving.serve_configs((2_000_000, 6_000_000))
    dstars = [wl.DeviceBatchStar(ctx, c) for c in cfgs]
    tmpl = serving.templates(64, seed=5)
    qs = [oc.BatchQuery.make(*t) for t in tmpl]
    plan = serving.arrivals(500.0, 300, len(dstars), len(qs), seed=2)
    srv = ctx.server(max_batch=64, streams=2, memory_budget=2 << 30)
    ids = [srv.add_dataset(d.star) for d in dstars]
    res = serving.run_open_loop(srv, ids, qs, plan, keep_results=True)
    _say({k: v for k, v in res.items() if k != "results"}, srv.stats())
    assert res["completed"] == len(plan) and res["failed"] == 0
    alone = ctx.server(max_batch=1, streams=1, memory_budget=2 << 30)
    aids = [alone.add_dataset(d.star) for d in dstars]
    for i, s, c in res["results"][::25]:
        a = plan[i]
        s1, c1 = alone.wait(alone.submit(aids[a.dataset], qs[a.template]))
        np.testing.assert_array_equal(s, s1, err_msg=f"arrival {i}")
        np.testing.assert_array_equal(c, c1, err_msg=f"arrival {i}")
    alone.close()
    srv.close()


def test_server_close_with_batches_in_flight(ctx):
    """Closing a server while batches are on the device or being copied out
    finishes (or drops, if not yet admitted) every request without hanging;
    the results waited for before the close are right, and a new server on
    the same datasets serves again."""
    cfgs = serving.serve_configs((20_000_000,))
    dstars = [wl.DeviceBatchStar(ctx, c) for c in cfgs]
    tmpl = serving.templates(64, seed=8)
    qs = [oc.BatchQuery.make(*t) for t in tmpl]
    srv = ctx.server(max_batch=64, streams=2, memory_budget=8 << 30)
    d = srv.add_dataset(dstars[0].star)
    tickets = [srv.submit(d, qs[i % len(qs)]) for i in range(192)]
    first = [srv.wait(t) for t in tickets[:4]]
    _say("closing with", len(tickets) - 4, "requests outstanding")
    srv.close()
    alone = ctx.server(max_batch=1, streams=1, memory_budget=8 << 30)
    a = alone.add_dataset(dstars[0].star)
    for i, (s, c) in enumerate(first):
        s1, c1 = alone.wait(alone.submit(a, qs[i]))
        np.testing.assert_array_equal(s, s1, err_msg=f"request {i}")
        np.testing.assert_array_equal(c, c1, err_msg=f"request {i}")
    alone.close()
