"""The serving loop on the CPU: admission policy, continuous batching and the
failure paths (budget rejection and queueing, allocation failure, executor
failure, timeout, cancellation), driven through olap_server_create_host with
a Python executor so no device is needed. The executor computes each query
with the oracle over small host star schemas."""
import threading
import time

import numpy as np
import pytest

from memory_database_amd import olapcore as oc
from memory_database_amd import workloads as wl
from oracle import pyoracle as po


def aligned(a, align=64):
    """A copy of `a` whose data starts on an `align`-byte boundary."""
    buf = np.empty(a.nbytes + align, dtype=np.uint8)
    off = (-buf.ctypes.data) % align
    out = buf[off:off + a.nbytes].view(a.dtype)
    out[:] = a
    return out


class HostStar:
    """A small batch star on the host (the device generator's streams)."""

    def __init__(self, rows, seed):
        cfg = wl.scaled(wl.STAR_4DIM_600M, rows, seed=seed)
        self.cfg = cfg
        self.filters, self.attrs, self.fks = [None] * 4, [None] * 4, [None] * 4
        for d in cfg.dims:
            self.filters[d.slot] = aligned(po.gen_uniform(d.rows, seed, 100 + d.slot, 0, 100))
            self.attrs[d.slot] = aligned(po.gen_uniform(d.rows, seed, 500 + d.slot, 0, wl.ATTR_HI))
            self.fks[d.slot] = aligned(po.gen_uniform(rows, seed, 300 + d.slot, 0, d.rows))
        self.m1 = aligned(po.gen_uniform(rows, seed, 400, cfg.m_lo, cfg.m_hi + 1))
        self.star = oc.Context.star(self.filters, self.attrs, self.fks, self.m1, None, rows=rows)

    def answer(self, q: oc.BatchQuery):
        d = q.as_dict()
        dvs = [None] * 4
        for (t, lo, hi, g) in d["steps"]:
            dvs[t] = po.build_dimvec(self.filters[t], self.attrs[t], lo, hi, 1 if g > 0 else 0, 4,
                                     group_mod=g)
        p = po.plan([s[0] for s in d["steps"]], d["factor"], d["group_nums"])
        s, c, bad = po.run("rowwise", p, po.Star(dvs, self.fks, self.m1), nthreads=2)
        assert bad == 0
        return s, c


def oracle_executor(stars, log=None, gate=None):
    def run(dataset, queries, offsets, osum, ocnt):
        if gate is not None:
            gate.wait()
        if log is not None:
            log.append((dataset, len(queries)))
        for i, q in enumerate(queries):
            s, c = stars[dataset].answer(q)
            osum[offsets[i]:offsets[i + 1]] = s
            ocnt[offsets[i]:offsets[i + 1]] = c
        return oc.OK
    return run


@pytest.fixture(scope="module")
def stars():
    return [HostStar(20_011, seed=101), HostStar(7_003, seed=202)]


def queries(n, seed, max_groups=500):
    return [oc.BatchQuery.make(*q) for q in wl.batch_queries(wl.STAR_4DIM_600M, n, seed, max_groups=max_groups)]


# ---------------------------------------------------------------- policy

def test_request_bytes_counts_dimension_vectors_and_group_table(stars):
    st = stars[0].star
    q = oc.BatchQuery.make([(2, 0, 10, 4), (0, 0, 50, 3)], [3, 1], 12)
    assert oc.server_request_bytes(st, q) == 16 * 12 + 4 * (200_000 + 30_000)
    q1 = oc.BatchQuery.make([(3, 0, 10, 0)], [1], 1)
    assert oc.server_request_bytes(st, q1) == 16 + 4 * 2_556


def test_plan_known_cases():
    # oldest dataset only, arrival order, max_batch
    assert oc.server_plan([1, 0, 1, 1, 1], [10] * 5, 1000, 3) == [0, 2, 3]
    # a request that does not fit is skipped, a later smaller one is taken
    assert oc.server_plan([0, 0, 0, 0], [40, 70, 20, 10], 100, 64) == [0, 2, 3]
    # the oldest waits when it does not fit, even if later ones would
    assert oc.server_plan([0, 0], [101, 1], 100, 64) == []
    assert oc.server_plan([], [], 100, 64) == []
    assert oc.server_plan([0], [100], 100, 64) == [0]


def test_plan_properties_random():
    rng = np.random.default_rng(5)
    for _ in range(500):
        n = int(rng.integers(1, 40))
        ds = rng.integers(0, 3, n)
        by = rng.integers(1, 200, n)
        budget = int(rng.integers(0, 1500))
        mb = int(rng.integers(1, 10))
        ch = oc.server_plan(ds, by, budget, mb)
        if by[0] > budget:
            assert ch == []
            continue
        assert ch and ch[0] == 0 and ch == sorted(ch) and len(ch) <= mb
        assert all(ds[i] == ds[0] for i in ch)
        assert sum(int(by[i]) for i in ch) <= budget
        # greedy: each same-dataset request passed over before the batch
        # filled did not fit what was left at its turn
        if len(ch) < mb:
            left = budget
            for i in range(n):
                if ds[i] != ds[0]:
                    continue
                if i in ch:
                    left -= int(by[i])
                else:
                    assert by[i] > left


# ---------------------------------------------------------------- the loop

def test_results_identical_alone_and_in_the_mix(stars):
    """Every request's result equals the oracle's, whether served alone
    (max_batch 1, one at a time) or in a mix (all submitted while paused,
    then admitted together over two datasets and two workers)."""
    qs = queries(40, seed=17)
    rng = np.random.default_rng(3)
    ds = rng.integers(0, 2, len(qs))
    expect = [stars[int(d)].answer(q) for d, q in zip(ds, qs)]

    alone = oc.Server.host(oracle_executor(stars), max_batch=1, streams=1)
    ids = [alone.add_dataset(s.star) for s in stars]
    for i, (d, q) in enumerate(zip(ds, qs)):
        s, c = alone.wait(alone.submit(ids[int(d)], q))
        np.testing.assert_array_equal(s, expect[i][0], err_msg=f"alone {i} sums")
        np.testing.assert_array_equal(c, expect[i][1], err_msg=f"alone {i} counts")
    assert alone.stats()["batches"] == len(qs)
    alone.close()

    log = []
    mix = oc.Server.host(oracle_executor(stars, log), max_batch=64, streams=2)
    ids = [mix.add_dataset(s.star) for s in stars]
    mix.pause(True)
    tickets = [mix.submit(ids[int(d)], q) for d, q in zip(ds, qs)]
    mix.pause(False)
    for i, t in enumerate(tickets):
        s, c = mix.wait(t)
        np.testing.assert_array_equal(s, expect[i][0], err_msg=f"mix {i} sums")
        np.testing.assert_array_equal(c, expect[i][1], err_msg=f"mix {i} counts")
    st = mix.stats()
    assert st["completed"] == len(qs) and st["failed"] == 0
    assert st["batches"] < len(qs) and st["mean_batch"] > 1
    assert sum(n for _, n in log) == len(qs)
    assert st["bytes_in_use"] == 0
    mix.close()


def test_budget_rejects_oversized_and_queues_the_rest(stars):
    st = stars[0].star
    q = oc.BatchQuery.make([(2, 0, 50, 100)], [1], 100)  # 1600 + 800000 bytes
    need = oc.server_request_bytes(st, q)
    gate = threading.Event()
    srv = oc.Server.host(oracle_executor(stars, gate=gate), max_batch=64, streams=2,
                         memory_budget=need + need // 2)
    d = srv.add_dataset(st)
    big = oc.BatchQuery.make([(2, 0, 50, 100), (0, 0, 50, 100)], [100, 1], 100_000)
    with pytest.raises(oc.OlapError) as e:
        srv.submit(d, oc.BatchQuery.make([(2, 0, 50, 100)], [1], 10_000_000))
    assert e.value.status == oc.E_CAPACITY
    assert oc.server_request_bytes(st, big) > need + need // 2
    with pytest.raises(oc.OlapError):
        srv.submit(d, big)
    t1 = srv.submit(d, q)
    deadline = time.time() + 10
    while srv.stats()["bytes_in_use"] != need and time.time() < deadline:
        time.sleep(0.005)
    assert srv.stats()["bytes_in_use"] == need      # t1 admitted (blocked in the executor)
    t2 = srv.submit(d, q)
    time.sleep(0.05)
    assert srv.stats()["bytes_in_use"] == need      # t2 queued: budget short
    with pytest.raises(oc.OlapError) as e:
        srv.wait(t2, timeout_us=2000)                # timeout while queued
    assert e.value.status == oc.E_BUSY
    gate.set()
    s1, c1 = srv.wait(t1)
    s2, c2 = srv.wait(t2)
    np.testing.assert_array_equal(s1, s2)
    np.testing.assert_array_equal(c1, c2)
    stt = srv.stats()
    assert stt["rejected"] == 2 and stt["completed"] == 2 and stt["bytes_in_use"] == 0
    assert stt["bytes_peak"] == need
    srv.close()


def test_cancel_pending_and_refuse_cancel_of_admitted(stars):
    gate = threading.Event()
    srv = oc.Server.host(oracle_executor([stars[1]], gate=gate), max_batch=1, streams=1)
    d = srv.add_dataset(stars[1].star)
    qs = queries(3, seed=4)
    t0 = srv.submit(d, qs[0])
    deadline = time.time() + 10
    while srv.stats()["bytes_in_use"] == 0 and time.time() < deadline:
        time.sleep(0.005)
    t1 = srv.submit(d, qs[1])
    t2 = srv.submit(d, qs[2])
    with pytest.raises(oc.OlapError) as e:
        srv.cancel(t0)                               # already in a batch
    assert e.value.status == oc.E_BUSY
    srv.cancel(t1)                                   # still pending
    with pytest.raises(oc.OlapError):
        srv.wait(t1)                                 # the ticket is gone
    gate.set()
    srv.wait(t0)
    s, c = srv.wait(t2)
    es, ec = stars[1].answer(qs[2])
    np.testing.assert_array_equal(s, es)
    np.testing.assert_array_equal(c, ec)
    st = srv.stats()
    assert st["cancelled"] == 1 and st["completed"] == 2
    srv.close()


def test_allocation_failure_fails_the_batch_and_the_server_recovers(stars):
    calls = {"n": 0}
    keep = {}

    def alloc(n):
        calls["n"] += 1
        if calls["n"] == 1:
            return 0                                 # the first staging allocation fails
        a = np.empty(n // 8 + 1, dtype=np.int64)
        keep[a.ctypes.data] = a
        return a.ctypes.data

    def free(p):
        keep.pop(p, None)

    srv = oc.Server.host(oracle_executor(stars), max_batch=64, streams=1)
    srv.set_allocator(alloc, free)
    d = srv.add_dataset(stars[0].star)
    qs = queries(2, seed=8)
    with pytest.raises(oc.OlapError) as e:
        srv.wait(srv.submit(d, qs[0]))
    assert e.value.status == oc.E_NOMEM
    s, c = srv.wait(srv.submit(d, qs[1]))
    es, ec = stars[0].answer(qs[1])
    np.testing.assert_array_equal(s, es)
    np.testing.assert_array_equal(c, ec)
    st = srv.stats()
    assert st["failed"] == 1 and st["completed"] == 1 and st["bytes_in_use"] == 0
    with pytest.raises(oc.OlapError):
        srv.set_allocator(alloc, free)               # only before the first submit
    srv.close()


def test_executor_failure_reaches_every_request_of_its_batch(stars):
    fail = {"on": True}

    def run(dataset, queries, offsets, osum, ocnt):
        if fail["on"]:
            return oc.E_HIP
        return oracle_executor(stars)(dataset, queries, offsets, osum, ocnt)

    srv = oc.Server.host(run, max_batch=64, streams=1)
    d = srv.add_dataset(stars[0].star)
    qs = queries(4, seed=12)
    srv.pause(True)
    ts = [srv.submit(d, q) for q in qs]
    srv.pause(False)
    for t in ts:
        with pytest.raises(oc.OlapError) as e:
            srv.wait(t)
        assert e.value.status == oc.E_HIP
    fail["on"] = False
    s, _ = srv.wait(srv.submit(d, qs[0]))
    np.testing.assert_array_equal(s, stars[0].answer(qs[0])[0])
    assert srv.stats()["failed"] == 4
    srv.close()


def test_invalid_requests_and_configs(stars):
    with pytest.raises(oc.OlapError):
        oc.Server.host(oracle_executor(stars), max_batch=65)
    with pytest.raises(oc.OlapError):
        oc.Server.host(oracle_executor(stars), streams=0)
    srv = oc.Server.host(oracle_executor(stars))
    d = srv.add_dataset(stars[0].star)
    with pytest.raises(oc.OlapError):
        srv.submit(d + 1, queries(1, seed=1)[0])     # unknown dataset
    with pytest.raises(oc.OlapError):
        srv.submit(d, oc.BatchQuery.make([(0, 0, 10, 5)], [1], 4))  # gid reaches G
    with pytest.raises(oc.OlapError):
        srv.submit(d, oc.BatchQuery.make([(1, 0, 10, 2), (1, 0, 10, 2)], [2, 1], 4))
    half = oc.Star()
    half.dim_rows[0] = 10
    half.filter[0] = 64
    half.attr[0] = 64
    half.rows = 0
    d2 = srv.add_dataset(half)
    with pytest.raises(oc.OlapError):                # table 1 is absent from that dataset
        srv.submit(d2, oc.BatchQuery.make([(1, 0, 10, 2)], [1], 2))
    srv.close()
