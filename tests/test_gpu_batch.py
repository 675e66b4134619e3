"""Batched form on the GPU: one fused scan for up to 64 queries, checked per
query against the oracle (each query restated as a plain star join over int32
dimension vectors), across autotuner configurations and hipGraph replays."""
import numpy as np
import pytest

from memory_database_amd import olapcore as oc
from memory_database_amd import workloads as wl
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def host_star(cfg, rows):
    """Host copy of the batch star (same generator streams as the device)."""
    filters, attrs, fks = [None] * 4, [None] * 4, [None] * 4
    for d in cfg.dims:
        filters[d.slot] = po.gen_uniform(d.rows, cfg.seed, 100 + d.slot, 0, 100)
        attrs[d.slot] = po.gen_uniform(d.rows, cfg.seed, 500 + d.slot, 0, wl.ATTR_HI)
        fks[d.slot] = po.gen_uniform(rows, cfg.seed, 300 + d.slot, 0, d.rows)
    m1 = po.gen_uniform(rows, cfg.seed, 400, cfg.m_lo, cfg.m_hi + 1)
    return filters, attrs, fks, m1


def oracle_query(host, steps, factor, G, ncpu):
    filters, attrs, fks, m1 = host
    dvs = [None] * 4
    for (t, lo, hi, g) in steps:
        dvs[t] = po.build_dimvec(filters[t], attrs[t], lo, hi, 1 if g > 0 else 0, 4, group_mod=g)
    p = po.plan([s[0] for s in steps], factor, G)
    s, c, bad = po.run("rowwise", p, po.Star(dvs, fks, m1), nthreads=ncpu)
    assert bad == 0
    return s, c


def make_batch(ctx, rows, nq, seed, max_groups=100_000):
    cfg = wl.scaled(wl.STAR_4DIM_600M, rows, name=f"batch_{rows}")
    ds = wl.DeviceBatchStar(ctx, cfg)
    qs = wl.batch_queries(cfg, nq, seed, max_groups=max_groups)
    bq = [oc.BatchQuery.make(*q) for q in qs]
    return cfg, ds, qs, ctx.batch(ds.star, bq)


def check_outputs(b, qs, host, osum, ocnt, ncpu, label):
    s = osum.cpu().numpy()
    c = ocnt.cpu().numpy()
    for i, (steps, factor, G) in enumerate(qs):
        es, ec = oracle_query(host, steps, factor, G, ncpu)
        lo, hi = b.offsets[i], b.offsets[i + 1]
        np.testing.assert_array_equal(c[lo:hi], ec, err_msg=f"{label} query {i} counts")
        np.testing.assert_array_equal(s[lo:hi], es, err_msg=f"{label} query {i} sums")


def test_batch_matches_oracle_every_config(ctx, ncpu):
    import torch

    rows = 1_000_003
    _say(f"batch of 64 queries over {rows} rows")
    cfg, ds, qs, b = make_batch(ctx, rows, 64, seed=11)
    host = host_star(cfg, rows)
    osum = torch.empty(b.total_slots, dtype=torch.int64, device="cuda")
    ocnt = torch.empty(b.total_slots, dtype=torch.int64, device="cuda")
    ref = None
    for k in range(ctx.lib.olap_batch_num_configs()):
        _say(f"  config {k}")
        b.set_config(k)
        b.run(osum, ocnt)
        torch.cuda.synchronize()
        got = (osum.cpu().numpy(), ocnt.cpu().numpy())
        if ref is None:
            check_outputs(b, qs, host, osum, ocnt, ncpu, f"config {k}")
            ref = got
        else:
            np.testing.assert_array_equal(got[0], ref[0], err_msg=f"config {k} sums")
            np.testing.assert_array_equal(got[1], ref[1], err_msg=f"config {k} counts")
    cfgk, th, rp, kib, staged, ms = b.autotune(osum, ocnt)
    _say(f"  autotune: config {cfgk} ({th} threads, {rp} rows/thread, {kib} KiB tables, "
         f"staged={staged}) {ms:.3f} ms")
    torch.cuda.synchronize()
    np.testing.assert_array_equal(osum.cpu().numpy(), ref[0])
    b.close()


@pytest.mark.parametrize("measures", ["two", "one_wide"])
def test_batch_two_measures_every_config(ctx, ncpu, measures):
    """Measures spanning the int32 range: two of them (SUM of m1 + m2,
    16-byte staged records), or one (6-byte one-pass records, the range being
    too wide for 4-byte ones). Every configuration equals the oracle."""
    import torch

    rows = 300_007
    cfg = wl.scaled(wl.STAR_4DIM_600M, rows, name="batch_m2")
    ds = wl.DeviceBatchStar(ctx, cfg)
    qs = wl.batch_queries(cfg, 24, 13, max_groups=30_000)
    host = list(host_star(cfg, rows))
    rng = np.random.default_rng(13)
    m1 = rng.integers(-2 ** 31, 2 ** 31 - 1, rows, dtype=np.int32)
    m2 = rng.integers(-2 ** 31, 2 ** 31 - 1, rows, dtype=np.int32) if measures == "two" else None
    dm1 = torch.from_numpy(m1).cuda()
    dm2 = torch.from_numpy(m2).cuda() if m2 is not None else None
    star = oc.Context.star(ds.filters, ds.attrs, ds.fks, dm1, dm2, rows=rows)
    b = ctx.batch(star, [oc.BatchQuery.make(*q) for q in qs])
    expect = []
    filters, attrs, fks, _ = host
    for steps, factor, G in qs:
        dvs = [None] * 4
        for (t, lo, hi, g) in steps:
            dvs[t] = po.build_dimvec(filters[t], attrs[t], lo, hi, 1 if g > 0 else 0, 4, group_mod=g)
        p = po.plan([st[0] for st in steps], factor, G)
        es, ec, bad = po.run("rowwise", p, po.Star(dvs, fks, m1, m2), nthreads=ncpu)
        assert bad == 0
        expect.append((es, ec))
    osum = torch.empty(b.total_slots, dtype=torch.int64, device="cuda")
    ocnt = torch.empty_like(osum)
    for k in range(ctx.lib.olap_batch_num_configs()):
        try:
            b.set_config(k)
        except oc.OlapError as e:
            _say(f"  config {k} skipped: {e}")
            continue
        b.run(osum, ocnt)
        torch.cuda.synchronize()
        s, c = osum.cpu().numpy(), ocnt.cpu().numpy()
        for i, (es, ec) in enumerate(expect):
            lo, hi = b.offsets[i], b.offsets[i + 1]
            np.testing.assert_array_equal(c[lo:hi], ec, err_msg=f"config {k} query {i} counts")
            np.testing.assert_array_equal(s[lo:hi], es, err_msg=f"config {k} query {i} sums")
    b.close()


def test_batch_filter_only_table_without_attribute_column(ctx, ncpu):
    """A table that the queries only filter on may come without a group
    attribute column (NULL); the small date table is then still copied into
    LDS by the configurations that keep small tables there. Every
    configuration equals the oracle."""
    import torch

    rows = 200_003
    cfg = wl.scaled(wl.STAR_4DIM_600M, rows, name="batch_noattr")
    ds = wl.DeviceBatchStar(ctx, cfg)
    host = list(host_star(cfg, rows))
    # date (slot 3) filter-only in every query that probes it
    qs = [([(3, 0, 50, 0), (0, 0, 30, 25)], [0, 1], 25),
          ([(3, 10, 60, 0)], [0], 1),
          ([(3, 0, 90, 0), (2, 0, 10, 7), (1, 0, 40, 5)], [0, 1, 7], 35),
          ([(1, 5, 70, 5), (3, 20, 100, 0)], [1, 0], 5),
          ([(0, 0, 100, 25)], [1], 25)]
    star = oc.Context.star(ds.filters, [ds.attrs[0], ds.attrs[1], ds.attrs[2], None], ds.fks, ds.m1, None,
                           rows=rows)
    b = ctx.batch(star, [oc.BatchQuery.make(*q) for q in qs])
    expect = [oracle_query(host, steps, factor, G, ncpu) for steps, factor, G in qs]
    osum = torch.empty(b.total_slots, dtype=torch.int64, device="cuda")
    ocnt = torch.empty_like(osum)
    for k in range(ctx.lib.olap_batch_num_configs()):
        try:
            b.set_config(k)
        except oc.OlapError as e:
            _say(f"  config {k} skipped: {e}")
            continue
        _say(f"  config {k}")
        b.run(osum, ocnt)
        torch.cuda.synchronize()
        s, c = osum.cpu().numpy(), ocnt.cpu().numpy()
        for i, (es, ec) in enumerate(expect):
            lo, hi = b.offsets[i], b.offsets[i + 1]
            np.testing.assert_array_equal(c[lo:hi], ec, err_msg=f"config {k} query {i} counts")
            np.testing.assert_array_equal(s[lo:hi], es, err_msg=f"config {k} query {i} sums")
    b.close()


@pytest.mark.parametrize("sizes", [(2_556, 30_000, 2_000, 2_000), (2_556, 30_000, 200_000, 60_000)],
                         ids=["one_table_in_hbm", "three_tables_in_hbm"])
def test_batch_tables_outside_lds_every_config(ctx, ncpu, sizes):
    """Stars whose configurations keep one or three of the four tables
    outside LDS (the kernels renumber those tables and spread their gathers
    by position). Every configuration equals the oracle."""
    import torch

    rows = 300_007
    d, c, p, s = sizes
    dims = [wl.Dim(3, d, 50, 7), wl.Dim(0, c, 20, 25), wl.Dim(2, p, 4, 1), wl.Dim(1, s, 20, 5)]
    cfg = wl.scaled(wl.STAR_4DIM_600M, rows, name=f"batch_sizes_{p}_{s}", dims=dims)
    ds = wl.DeviceBatchStar(ctx, cfg)
    qs = wl.batch_queries(cfg, 24, 17, max_groups=20_000)
    host = host_star(cfg, rows)
    b = ctx.batch(ds.star, [oc.BatchQuery.make(*q) for q in qs])
    expect = [oracle_query(host, steps, factor, G, ncpu) for steps, factor, G in qs]
    osum = torch.empty(b.total_slots, dtype=torch.int64, device="cuda")
    ocnt = torch.empty_like(osum)
    for k in range(ctx.lib.olap_batch_num_configs()):
        try:
            b.set_config(k)
        except oc.OlapError as e:
            _say(f"  config {k} skipped: {e}")
            continue
        _say(f"  config {k}")
        b.run(osum, ocnt)
        torch.cuda.synchronize()
        sm, cn = osum.cpu().numpy(), ocnt.cpu().numpy()
        for i, (es, ec) in enumerate(expect):
            lo, hi = b.offsets[i], b.offsets[i + 1]
            np.testing.assert_array_equal(cn[lo:hi], ec, err_msg=f"config {k} query {i} counts")
            np.testing.assert_array_equal(sm[lo:hi], es, err_msg=f"config {k} query {i} sums")
    b.close()


def test_batch_small_groups_and_ragged_rows(ctx, ncpu):
    """Small group counts (LDS tables for every query), a ragged row count and
    a batch of one."""
    import torch

    for rows, nq, maxg in [(8191, 17, 50), (5, 1, 3), (300_001, 64, 2000)]:
        _say(f"batch rows={rows} nq={nq} max_groups={maxg}")
        cfg, ds, qs, b = make_batch(ctx, rows, nq, seed=rows, max_groups=maxg)
        host = host_star(cfg, rows)
        osum = torch.empty(b.total_slots, dtype=torch.int64, device="cuda")
        ocnt = torch.empty(b.total_slots, dtype=torch.int64, device="cuda")
        b.run(osum, ocnt)
        torch.cuda.synchronize()
        check_outputs(b, qs, host, osum, ocnt, ncpu, f"rows={rows}")
        b.close()


def test_batch_equals_plain_queries(ctx):
    """Each query of the batch equals the plain star join of the same query."""
    import torch

    rows = 2_000_000
    cfg, ds, qs, b = make_batch(ctx, rows, 24, seed=5)
    osum = torch.empty(b.total_slots, dtype=torch.int64, device="cuda")
    ocnt = torch.empty(b.total_slots, dtype=torch.int64, device="cuda")
    b.run(osum, ocnt)
    torch.cuda.synchronize()
    for i, (steps, factor, G) in enumerate(qs):
        plan, cols = ds.plain_query(steps, factor, G)
        ps = torch.empty(G, dtype=torch.int64, device="cuda")
        pc = torch.empty(G, dtype=torch.int64, device="cuda")
        q = ctx.query(plan, cols)
        q.run("rowwise", ps, pc)
        torch.cuda.synchronize()
        lo, hi = b.offsets[i], b.offsets[i + 1]
        assert torch.equal(ps, osum[lo:hi]) and torch.equal(pc, ocnt[lo:hi]), f"query {i}"
        q.close()
    b.close()


def test_batch_graph_replays(ctx):
    """A captured graph replayed several times, with eager work and
    synchronisations between replays: every replay equals the eager run."""
    import torch

    rows = 3_000_000
    cfg, ds, qs, b = make_batch(ctx, rows, 64, seed=9)
    osum = torch.empty(b.total_slots, dtype=torch.int64, device="cuda")
    ocnt = torch.empty(b.total_slots, dtype=torch.int64, device="cuda")
    b.run(osum, ocnt)
    torch.cuda.synchronize()
    es, ec = osum.clone(), ocnt.clone()
    stream = torch.cuda.Stream()
    b.capture(osum, ocnt, stream.cuda_stream)
    for rep in range(4):
        osum.fill_(-1)
        ocnt.fill_(-1)
        junk = torch.randn(1 << 20, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        b.launch(stream.cuda_stream)
        torch.cuda.synchronize()
        assert torch.equal(osum, es) and torch.equal(ocnt, ec), f"replay {rep}"
        del junk
    b.close()


def test_batch_rejects_bad_queries(ctx):
    import torch

    cfg, ds, qs, b = make_batch(ctx, 1000, 2, seed=1)
    b.close()
    bad = oc.BatchQuery.make([(0, 0, 10, 5), (0, 0, 10, 5)], [5, 1], 25)  # table twice
    with pytest.raises(oc.OlapError):
        ctx.batch(ds.star, [bad])
    bad = oc.BatchQuery.make([(0, 0, 10, 5)], [1], 4)  # gid 4 reaches group_nums
    with pytest.raises(oc.OlapError):
        ctx.batch(ds.star, [bad])
    with pytest.raises(oc.OlapError):
        ctx.batch(ds.star, [oc.BatchQuery.make([(1, 0, 10, 2)], [1], 2)] * 65)


@pytest.mark.parametrize("aggregation,new_measures", [(1, False), (2, False), (2, True)])
def test_batch_graph_replay_after_columns_grow_matches(ctx, aggregation, new_measures):
    """A graph captured with record buffers sized for few matches, replayed
    after the fact columns change so that every row matches: the matches that
    no longer fit the buffers go to the outputs directly, and the replay
    equals a fresh batch over the new columns and the closed form. With
    new_measures the measures also leave the range the one-pass records were
    packed for (4-byte records take a measure offset of 20 bits), so those
    matches are added directly too."""
    import torch

    rows, D, G = 1_000_000, 50_000, 20_000
    filt = torch.arange(D, dtype=torch.int32, device="cuda") % 100        # filter 0: 1 row in 100
    attr = torch.arange(D, dtype=torch.int32, device="cuda") * 7 % 65536
    fk = torch.full((rows,), 99, dtype=torch.int32, device="cuda")         # row 99 fails [0, 1)
    fk[:1000] = torch.arange(1000, dtype=torch.int32, device="cuda") * 100 % D
    gen = torch.Generator(device="cuda").manual_seed(5)
    m1 = torch.randint(-1000, 1000, (rows,), dtype=torch.int32, device="cuda", generator=gen)
    star = oc.Context.star([filt, None, None, None], [attr, None, None, None],
                           [fk, None, None, None], m1, None, rows=rows)
    q = oc.BatchQuery.make([(0, 0, 1, G)], [1], G)
    cfg = next(k for k in range(ctx.lib.olap_batch_num_configs())
               if oc.config_info(k)[3] == aggregation and oc.config_info(k)[2] == 8)
    _say(f"config {cfg} {oc.config_info(cfg)}")
    b = ctx.batch(star, [q])
    b.set_config(cfg)
    osum = torch.empty(b.total_slots, dtype=torch.int64, device="cuda")
    ocnt = torch.empty_like(osum)
    b.run(osum, ocnt)
    torch.cuda.synchronize()
    assert int(ocnt.sum()) == 1000
    stream = torch.cuda.Stream()
    b.capture(osum, ocnt, stream.cuda_stream)
    # every row now probes a passing dimension row
    fk.copy_(torch.randint(0, D // 100, (rows,), dtype=torch.int32, device="cuda", generator=gen) * 100)
    if new_measures:
        m1.copy_(torch.randint(-2 ** 30, 2 ** 30, (rows,), dtype=torch.int32, device="cuda", generator=gen))
    torch.cuda.synchronize()
    b.launch(stream.cuda_stream)
    torch.cuda.synchronize()
    gid = (attr[fk.long()].long() % G)
    es = torch.zeros(G, dtype=torch.int64, device="cuda").index_add_(0, gid, m1.long())
    ec = torch.bincount(gid, minlength=G)
    assert int(ocnt.sum()) == rows
    assert torch.equal(ocnt, ec) and torch.equal(osum, es), "replay after the columns changed"
    b2 = ctx.batch(star, [q])
    b2.set_config(cfg)
    fsum, fcnt = torch.empty_like(osum), torch.empty_like(ocnt)
    b2.run(fsum, fcnt)
    torch.cuda.synchronize()
    assert torch.equal(fsum, es) and torch.equal(fcnt, ec), "fresh batch"
    b2.close()
    b.close()


def test_batch_segments_overflow_into_residual_passes(ctx, ncpu):
    """Fact keys concentrated on the dimension rows that every query passes.
    The segmented kernel sizes each query's event segment from the pass
    fractions of the dimension rows (2% per table here), but most fact rows
    match most queries, so the segments overflow in every tile and the
    residual passes place the rest. Every configuration equals the oracle."""
    import torch

    rows = 200_003
    rng = np.random.default_rng(5)
    dims = [2556, 30000, 200000, 2000]
    filters = [rng.integers(0, 100, d, dtype=np.int32) for d in dims]
    attrs = [rng.integers(0, 1 << 20, d, dtype=np.int32) for d in dims]
    fks = []
    for t, d in enumerate(dims):
        hot = np.flatnonzero(filters[t] < 2)
        pick_hot = rng.random(rows) < 0.8
        fks.append(np.where(pick_hot, rng.choice(hot, rows), rng.integers(0, d, rows)).astype(np.int32))
    m1 = rng.integers(1000, 1_000_001, rows, dtype=np.int32)
    qs = []
    for q in range(64):
        k = int(rng.integers(1, 5))
        tables = [int(x) for x in rng.permutation(4)[:k]]
        gs = [int(rng.integers(1, 40)) for _ in tables]
        steps = [(t, 0, 2, g) for t, g in zip(tables, gs)]
        factor = [int(np.prod(gs[i + 1:], dtype=np.int64)) for i in range(k)]
        qs.append((steps, factor, int(np.prod(gs, dtype=np.int64))))
    dev = lambda a: torch.from_numpy(a).cuda()
    star = oc.Context.star([dev(f) for f in filters], [dev(a) for a in attrs], [dev(f) for f in fks], dev(m1),
                           None, rows=rows)
    b = ctx.batch(star, [oc.BatchQuery.make(*q) for q in qs])
    expect = []
    for steps, factor, G in qs:
        dvs = [None] * 4
        for (t, lo, hi, g) in steps:
            dvs[t] = po.build_dimvec(filters[t], attrs[t], lo, hi, 1 if g > 0 else 0, 4, group_mod=g)
        p = po.plan([st[0] for st in steps], factor, G)
        es, ec, bad = po.run("rowwise", p, po.Star(dvs, fks, m1), nthreads=ncpu)
        assert bad == 0
        expect.append((es, ec))
    _say(f"  {sum(int(ec.sum()) for _, ec in expect) / rows:.1f} matches per fact row "
         f"(expected from the pass fractions: {b.expected_events():.4f})")
    osum = torch.empty(b.total_slots, dtype=torch.int64, device="cuda")
    ocnt = torch.empty_like(osum)
    ran = 0
    for k in range(ctx.lib.olap_batch_num_configs()):
        try:
            b.set_config(k)
        except oc.OlapError as e:
            assert e.status == oc.E_CAPACITY, e
            continue
        b.run(osum, ocnt)
        torch.cuda.synchronize()
        s, c = osum.cpu().numpy(), ocnt.cpu().numpy()
        for i, (es, ec) in enumerate(expect):
            lo, hi = b.offsets[i], b.offsets[i + 1]
            np.testing.assert_array_equal(c[lo:hi], ec, err_msg=f"config {k} query {i} counts")
            np.testing.assert_array_equal(s[lo:hi], es, err_msg=f"config {k} query {i} sums")
        ran += 1
    assert ran > 0
    b.close()


@pytest.mark.parametrize("attrs", ["below_2_31", "negative"])
def test_batch_group_terms_wide_attributes_and_moduli(ctx, ncpu, attrs):
    """Group-by terms at the edges of the modulo arithmetic: attributes over
    the whole non-negative int32 range (the 32-bit multiply-high modulo) or
    with negative values (read as 2^31 and above: the 64-bit one), and
    moduli at and around powers of two up to 2^24 - 1. Every configuration
    equals the oracle."""
    import torch

    rows = 300_007
    rng = np.random.default_rng(17)
    dims = [2556, 30000, 200000, 2000]
    filters = [rng.integers(0, 100, d, dtype=np.int32) for d in dims]
    lo_attr = -2 ** 31 if attrs == "negative" else 0
    attr_cols = [rng.integers(lo_attr, 2 ** 31, d, dtype=np.int64).astype(np.int32) for d in dims]
    for a in attr_cols:   # the extremes of the range
        a[:4] = [0, 1, 2 ** 31 - 1, 2 ** 31 - 2] if attrs != "negative" else [-1, -2 ** 31, 2 ** 31 - 1, 0]
    fks = [rng.integers(0, d, rows, dtype=np.int32) for d in dims]
    m1 = rng.integers(1000, 1_000_001, rows, dtype=np.int32)
    big = [2, 3, 4, 7, 8, 255, 256, 257, 65535, 65536, 65537, 99991, (1 << 20) + 1, (1 << 23) - 1,
           (1 << 24) - 1]
    qs = []
    for i, g in enumerate(big):   # one table grouped by a large modulo
        t = i % 4
        qs.append(([(t, 0, 60, g)], [1], g))
    for i in range(14):           # two tables
        t1, t2 = [int(x) for x in rng.permutation(4)[:2]]
        g1 = int(rng.choice([2, 3, 5, 16, 17, 255, 256, 257, 4095, 4096]))
        g2 = int(rng.integers(1, 4097))
        qs.append(([(t1, 10, 70, g1), (t2, 0, 50, g2)], [g2, 1], g1 * g2))
    dev = lambda a: torch.from_numpy(a).cuda()
    star = oc.Context.star([dev(f) for f in filters], [dev(a) for a in attr_cols], [dev(f) for f in fks], dev(m1),
                           None, rows=rows)
    b = ctx.batch(star, [oc.BatchQuery.make(*q) for q in qs])
    expect = []
    for steps, factor, G in qs:
        dvs = [None] * 4
        for (t, lo, hi, g) in steps:
            dvs[t] = po.build_dimvec(filters[t], attr_cols[t], lo, hi, 1 if g > 0 else 0, 4, group_mod=g)
        p = po.plan([st[0] for st in steps], factor, G)
        es, ec, bad = po.run("rowwise", p, po.Star(dvs, fks, m1), nthreads=ncpu)
        assert bad == 0
        expect.append((es, ec))
    osum = torch.empty(b.total_slots, dtype=torch.int64, device="cuda")
    ocnt = torch.empty_like(osum)
    ran = 0
    for k in range(ctx.lib.olap_batch_num_configs()):
        try:
            b.set_config(k)
        except oc.OlapError as e:
            _say(f"  config {k} skipped: {e}")
            continue
        b.run(osum, ocnt)
        torch.cuda.synchronize()
        s, c = osum.cpu().numpy(), ocnt.cpu().numpy()
        for i, (es, ec) in enumerate(expect):
            lo, hi = b.offsets[i], b.offsets[i + 1]
            np.testing.assert_array_equal(c[lo:hi], ec, err_msg=f"config {k} query {i} counts")
            np.testing.assert_array_equal(s[lo:hi], es, err_msg=f"config {k} query {i} sums")
        ran += 1
    assert ran > 0
    b.close()
