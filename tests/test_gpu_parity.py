"""GPU parity: every strategy of libolapcore (through the C ABI) against the
CPU oracle on the same inputs, bit-exact (int64 sums and counts per slot)."""
import numpy as np
import pytest

from memory_database_amd import olapcore as oc
from memory_database_amd import workloads as wl
from oracle import pyoracle as po
from tests.cpu_star import cpu_star

pytestmark = pytest.mark.gpu

STRATS = ["rowwise", "vwm_dv", "vwm_sv", "cwm_dv", "cwm_sv"]


def _say(*a):
    print(*a, flush=True)


def _dev(a, torch):
    return None if a is None else torch.from_numpy(np.ascontiguousarray(a)).cuda()


def gpu_run_all(ctx, plan_np, dimvecs, fks, m1, m2, strategies=STRATS, offset=0):
    """Upload numpy columns and run each strategy; offset > 0 shifts every
    fact column by that many elements (unaligned 128-bit path)."""
    import torch

    def up(a):
        if a is None:
            return None
        if offset:
            buf = torch.empty(a.size + offset, dtype=torch.from_numpy(a[:0]).dtype, device="cuda")
            buf[offset:] = torch.from_numpy(np.ascontiguousarray(a)).cuda()
            return buf[offset:]
        return _dev(a, torch)

    dv = [_dev(d, torch) for d in dimvecs]
    fk = [up(f) for f in fks]
    g1, g2 = up(m1), up(m2)
    rows = int(m1.size)
    plan = oc.plan_from(list(plan_np.orders)[: plan_np.dimvec_nums],
                        list(plan_np.factor)[: plan_np.dimvec_nums], int(plan_np.group_nums))
    cols = oc.Context.columns(dv, fk, g1, g2, rows=rows)
    G = int(plan.group_nums)
    out = {}
    q = ctx.query(plan, cols)
    for s in strategies:
        _say(f"  strategy {s} rows={rows} G={G} offset={offset}")
        osum = torch.full((G,), -7, dtype=torch.int64, device="cuda")
        ocnt = torch.full((G,), -7, dtype=torch.int64, device="cuda")
        q.run(s, osum, ocnt)
        torch.cuda.synchronize()
        out[s] = (osum.cpu().numpy(), ocnt.cpu().numpy())
    q.close()
    return out


def check_all(ctx, plan_np, dimvecs, fks, m1, m2, ncpu, offset=0, strategies=STRATS):
    star = po.Star(dimvecs, fks, m1, m2)
    es, ec, bad = po.run("rowwise", plan_np, star, nthreads=ncpu)
    assert bad == 0
    got = gpu_run_all(ctx, plan_np, dimvecs, fks, m1, m2, strategies, offset)
    for s, (gs, gc) in got.items():
        np.testing.assert_array_equal(gc, ec, err_msg=f"{s}: counts")
        np.testing.assert_array_equal(gs, es, err_msg=f"{s}: sums")
    return es, ec


def random_star(rng, ndims, rows, code_bytes=1, max_groups=20, oob=True, big_dim=False,
                two_measures=None, wide_measure=True):
    slots = list(rng.permutation(4)[:ndims])
    groups = [0, 0, 0, 0]
    bitmap = [0, 0, 0, 0]
    sele = [0.0, 0.0, 0.0, 0.0]
    dimvecs = [None] * 4
    fks = [None] * 4
    for t in slots:
        D = int(rng.integers(1, 300_000 if big_dim else 5000))
        g = int(rng.integers(1, max_groups + 1))
        bm = int(rng.integers(0, 2)) if g == 1 else 1
        groups[t], bitmap[t] = g, bm
        thr = int(rng.integers(0, 101))
        sele[t] = thr / 100.0
        f = rng.integers(0, 100, D, dtype=np.int32)
        gc = rng.integers(0, g, D, dtype=np.int32)
        dimvecs[t] = po.build_dimvec(f, gc, -(2 ** 31), thr, bm, code_bytes)
        hi = D + (D // 8 + 2 if oob else 0)
        fks[t] = rng.integers(-3 if oob else 0, hi, rows, dtype=np.int32)
    plan = po.make_plan(groups, bitmap, sele)
    if wide_measure:
        m1 = rng.integers(-2 ** 31, 2 ** 31 - 1, rows, dtype=np.int32)
    else:
        m1 = rng.integers(1, 100, rows, dtype=np.int32)
    if two_measures is None:
        two_measures = bool(rng.integers(0, 2))
    m2 = rng.integers(-2 ** 31, 2 ** 31 - 1, rows, dtype=np.int32) if two_measures else None
    return plan, dimvecs, fks, m1, m2


def test_generators_match_oracle(ctx):
    import torch

    _say("generators: uniform + zipf, 1M rows")
    n = 1_000_003
    t = torch.empty(n, dtype=torch.int32, device="cuda")
    ctx.gen_uniform(t, 77, 5, -1000, 123457, first_row=11)
    np.testing.assert_array_equal(t.cpu().numpy(), po.gen_uniform(n, 77, 5, -1000, 123457, 11))
    thr = oc.zipf_table(200_000, 0.99)
    np.testing.assert_array_equal(thr, po.zipf_table(200_000, 0.99))
    thr_d = torch.from_numpy(thr.view(np.int64)).cuda()
    ctx.gen_zipf(t, 78, 6, thr_d, first_row=3)
    np.testing.assert_array_equal(t.cpu().numpy(), po.gen_zipf(n, 78, 6, thr, 3))


def test_build_dimvec_matches_oracle(ctx):
    import torch

    rng = np.random.default_rng(5)
    for code_bytes, bitmap, mod in [(1, 1, 126), (1, 0, 0), (4, 1, 0), (4, 1, 1000)]:
        _say(f"build_dimvec code_bytes={code_bytes} bitmap={bitmap} mod={mod}")
        D = 77_777
        f = rng.integers(-50, 150, D, dtype=np.int32)
        g = rng.integers(0, 100_000, D, dtype=np.int32)
        if code_bytes == 1:
            g = g % 126
        lo, hi = -3, 61
        out = torch.empty(D, dtype=torch.int8 if code_bytes == 1 else torch.int32, device="cuda")
        ctx.build_dimvec(_dev(f, torch), _dev(g, torch), lo, hi, bitmap, out, group_mod=mod)
        exp = po.build_dimvec(f, g, lo, hi, bitmap, code_bytes, group_mod=mod)
        np.testing.assert_array_equal(out.cpu().numpy(), exp)


@pytest.mark.parametrize("seed", list(range(12)))
def test_random_small_stars(ctx, ncpu, seed):
    rng = np.random.default_rng(1000 + seed)
    ndims = 1 + seed % 4
    rows = [0, 1, 3, 5, 4097, 8191, 8193, 20_000, 65_536 + 7, 100_000, 33, 250_001][seed]
    _say(f"random star seed={seed} ndims={ndims} rows={rows}")
    plan, dv, fk, m1, m2 = random_star(rng, ndims, rows, code_bytes=1 if seed % 3 else 4)
    check_all(ctx, plan, dv, fk, m1, m2, ncpu, offset=1 if seed % 5 == 4 else 0)


def test_wide_codes_and_global_group_table(ctx, ncpu):
    """int32 codes, 40,000 groups: the group table does not fit in LDS."""
    rng = np.random.default_rng(9)
    _say("wide codes / global table")
    plan, dv, fk, m1, m2 = random_star(rng, 2, 300_000, code_bytes=4, max_groups=200)
    assert plan.group_nums > 10_000 or True
    check_all(ctx, plan, dv, fk, m1, m2, ncpu)
    # single dimension with 100,000 groups
    D = 150_000
    f = rng.integers(0, 100, D, dtype=np.int32)
    g = rng.integers(0, 100_000, D, dtype=np.int32)
    dvec = po.build_dimvec(f, g, 0, 60, 1, 4)
    fks = [rng.integers(0, D, 500_000, dtype=np.int32), None, None, None]
    dvs = [dvec, None, None, None]
    p = po.plan([0], [1], 100_000)
    m = rng.integers(0, 1000, 500_000, dtype=np.int32)
    check_all(ctx, p, dvs, fks, m, None, ncpu)


@pytest.mark.parametrize("G,sel,rows,two,skew", [(100_000, 60, 3_000_001, False, False),
                                                 (1_000_000, 90, 2_000_000, True, False),
                                                 (1_000_000, 90, 2_000_000, False, True),
                                                 (5_000, 100, 1_000_003, False, False)])
def test_staged_aggregation_matches_oracle(ctx, ncpu, G, sel, rows, two, skew):
    """Group tables beyond LDS: global atomics, two-pass staged (count /
    scatter / reduce), global atomics behind the LDS cache of hot groups,
    one-pass staged (records into pool chunks, reduce) and the automatic
    choice all give the oracle's result, with uniform group codes and with
    codes skewed onto a few hot groups."""
    import torch

    rng = np.random.default_rng(G + sel)
    _say(f"staged aggregation G={G} sel={sel}% rows={rows} two_measures={two} skew={skew}")
    D = 400_000
    f = rng.integers(0, 100, D, dtype=np.int32)
    if skew:  # Zipf-like: most dimension rows carry one of a few group codes
        g = np.minimum(rng.zipf(1.3, D) - 1, G - 1).astype(np.int32)
    else:
        g = rng.integers(0, G, D, dtype=np.int32)
    dv = po.build_dimvec(f, g, 0, sel, 1, 4)
    D2 = 3000
    dv2 = po.build_dimvec(rng.integers(0, 100, D2, dtype=np.int32), np.zeros(D2, np.int32),
                          0, 95, 0, 1)
    dvs = [dv, None, dv2, None]
    fks = [rng.integers(0, D, rows, dtype=np.int32), None,
           rng.integers(0, D2 + 5, rows, dtype=np.int32), None]
    m1 = rng.integers(-2 ** 31, 2 ** 31 - 1, rows, dtype=np.int32)
    m2 = rng.integers(0, 1000, rows, dtype=np.int32) if two else None
    p = po.plan([2, 0], [0, 1], G)
    es, ec, _ = po.run("rowwise", p, po.Star(dvs, fks, m1, m2), nthreads=ncpu)
    plan = oc.plan_from([2, 0], [0, 1], G)
    cols = oc.Context.columns([_dev(d, torch) for d in dvs], [_dev(x, torch) for x in fks],
                              _dev(m1, torch), _dev(m2, torch), rows=rows)
    q = ctx.query(plan, cols)
    osum = torch.empty(G, dtype=torch.int64, device="cuda")
    ocnt = torch.empty(G, dtype=torch.int64, device="cuda")
    for layout in (0, 1, 2, 3, -1, 1, 2, 3):
        _say(f"  layout {layout}")
        q.set_aggregation(layout)
        q.run("rowwise", osum, ocnt)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(ocnt.cpu().numpy(), ec, err_msg=f"layout {layout}")
        np.testing.assert_array_equal(osum.cpu().numpy(), es, err_msg=f"layout {layout}")
    q.set_aggregation(-1)
    shape, threads, rpt, ms = q.autotune("rowwise", osum, ocnt)
    _say(f"  autotune -> {shape} ({ms:.3f} ms)")
    np.testing.assert_array_equal(osum.cpu().numpy(), es)
    q.close()


@pytest.mark.parametrize("case", ["streamed_second", "one_streamed", "count_only", "pool_overflow",
                                  "tail_only", "one_dim"])
def test_onepass_matches_oracle(ctx, ncpu, case):
    """One-pass staged aggregation (layout 3): with the first step passing
    most rows and almost none, counting only, a
    pool sized from a pass estimate far below the real matches (records past
    the pool are added directly), fewer rows than one tile (tail loop only),
    and a single dimension; every slot equals the oracle's."""
    import torch

    rng = np.random.default_rng(sum(map(ord, case)))
    G0 = 100_000  # groups of the first table; the second multiplies them by 3
    rows = {"tail_only": 5_001, "pool_overflow": 1_500_000}.get(case, 2_000_003)
    D = 50_000
    thr = {"one_streamed": 1, "pool_overflow": 1}.get(case, 70)
    f = rng.integers(0, 100, D, dtype=np.int32)
    g = rng.integers(0, G0, D, dtype=np.int32)
    dv = po.build_dimvec(f, g, 0, thr, 1, 4)
    if case == "pool_overflow":  # every fact row finds a passing dimension row
        passing = np.flatnonzero(f < thr).astype(np.int32)
        fk0 = passing[rng.integers(0, passing.size, rows)]
    else:
        fk0 = rng.integers(0, D, rows, dtype=np.int32)
    if case == "one_dim":
        dvs, fks, orders, factor, G = [dv, None, None, None], [fk0, None, None, None], [0], [1], G0
    else:
        D2 = 2000
        dv2 = po.build_dimvec(rng.integers(0, 100, D2, dtype=np.int32),
                              rng.integers(0, 3, D2, dtype=np.int32), 0, 80, 1, 4)
        dvs = [dv, None, None, dv2]
        fks = [fk0, None, None, rng.integers(0, D2, rows, dtype=np.int32)]
        orders, factor, G = [0, 3], [1, G0], 3 * G0
    m1 = None if case == "count_only" else rng.integers(-2 ** 31, 2 ** 31 - 1, rows, dtype=np.int32)
    m2 = rng.integers(0, 1000, rows, dtype=np.int32) if case == "streamed_second" else None
    _say(f"one-pass staged {case}: rows={rows} G={G}")
    p = po.plan(orders, factor, G)
    m1_oracle = np.zeros(rows, np.int32) if m1 is None else m1  # the oracle always reads a measure
    es, ec, bad = po.run("rowwise", p, po.Star(dvs, fks, m1_oracle, m2), nthreads=ncpu)
    assert bad == 0
    plan = oc.plan_from(orders, factor, G)
    cols = oc.Context.columns([_dev(d, torch) for d in dvs], [_dev(x, torch) for x in fks],
                              _dev(m1, torch), _dev(m2, torch), rows=rows)
    q = ctx.query(plan, cols)
    osum = torch.full((G,), -7, dtype=torch.int64, device="cuda")
    ocnt = torch.full((G,), -7, dtype=torch.int64, device="cuda")
    q.set_aggregation(3)
    for rep in range(2):
        q.run("rowwise", osum if m1 is not None else None, ocnt)
        torch.cuda.synchronize()
        th, rp, agg = q.launch_info()
        assert agg == "one-pass staged"
        np.testing.assert_array_equal(ocnt.cpu().numpy(), ec, err_msg=f"{case} counts, run {rep}")
        if m1 is not None:
            np.testing.assert_array_equal(osum.cpu().numpy(), es, err_msg=f"{case} sums, run {rep}")
    q.close()


def test_staged_forms_refuse_partitions_beyond_lds(ctx, ncpu):
    """2^27 groups make 16,384 partitions, whose per-partition state does not
    fit LDS: both staged forms are refused with E_CAPACITY, and the automatic
    choice (few expected matches: global atomics) still equals the oracle."""
    import torch

    G = 1 << 27
    rows = 200_003
    rng = np.random.default_rng(27)
    _say(f"G=2^27 over {rows} rows")
    D = 100_000
    dv = po.build_dimvec(rng.integers(0, 100, D, dtype=np.int32), rng.integers(0, G, D, dtype=np.int32),
                         0, 50, 1, 4)
    dvs = [dv, None, None, None]
    fks = [rng.integers(0, D, rows, dtype=np.int32), None, None, None]
    m1 = rng.integers(1, 1000, rows, dtype=np.int32)
    p = po.plan([0], [1], G)
    es, ec, _ = po.run("rowwise", p, po.Star(dvs, fks, m1), nthreads=ncpu)
    plan = oc.plan_from([0], [1], G)
    cols = oc.Context.columns([_dev(d, torch) for d in dvs], [_dev(x, torch) for x in fks],
                              _dev(m1, torch), None, rows=rows)
    q = ctx.query(plan, cols)
    osum = torch.empty(G, dtype=torch.int64, device="cuda")
    ocnt = torch.empty(G, dtype=torch.int64, device="cuda")
    for layout in (1, 3):
        q.set_aggregation(layout)
        with pytest.raises(oc.OlapError) as err:
            q.run("rowwise", osum, ocnt)
        assert err.value.status == oc.E_CAPACITY
    q.set_aggregation(-1)
    q.run("rowwise", osum, ocnt)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(ocnt.cpu().numpy(), ec)
    np.testing.assert_array_equal(osum.cpu().numpy(), es)
    # the autotuner skips the refused forms and leaves a complete run
    shape, threads, rpt, ms = q.autotune("rowwise", osum, ocnt)
    _say(f"  autotune -> {shape} ({ms:.3f} ms)")
    np.testing.assert_array_equal(ocnt.cpu().numpy(), ec)
    np.testing.assert_array_equal(osum.cpu().numpy(), es)
    q.close()


def test_automatic_layout_beyond_partition_state_takes_atomics(ctx, ncpu):
    """2^27 groups with ~2.4M expected matches (4M rows, 60% pass): the
    automatic choice would stage, but both staged forms are refused for
    partition state beyond LDS, so it takes global atomics and still equals
    the oracle."""
    import torch

    G = 1 << 27
    rows = 4_000_000
    rng = np.random.default_rng(127)
    _say(f"G=2^27, {rows} rows, automatic layout")
    D = 100_000
    dv = po.build_dimvec(rng.integers(0, 100, D, dtype=np.int32), rng.integers(0, G, D, dtype=np.int32),
                         0, 60, 1, 4)
    dvs = [dv, None, None, None]
    fks = [rng.integers(0, D, rows, dtype=np.int32), None, None, None]
    m1 = rng.integers(1, 1000, rows, dtype=np.int32)
    p = po.plan([0], [1], G)
    es, ec, _ = po.run("rowwise", p, po.Star(dvs, fks, m1), nthreads=ncpu)
    assert ec.sum() > 2_000_000
    plan = oc.plan_from([0], [1], G)
    cols = oc.Context.columns([_dev(d, torch) for d in dvs], [_dev(x, torch) for x in fks],
                              _dev(m1, torch), None, rows=rows)
    q = ctx.query(plan, cols)
    osum = torch.empty(G, dtype=torch.int64, device="cuda")
    ocnt = torch.empty(G, dtype=torch.int64, device="cuda")
    q.set_aggregation(-1)
    q.run("rowwise", osum, ocnt)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(ocnt.cpu().numpy(), ec)
    np.testing.assert_array_equal(osum.cpu().numpy(), es)
    q.close()


def test_large_dimensions_spill_to_hbm(ctx, ncpu):
    """Dimension vectors larger than LDS are probed from HBM."""
    rng = np.random.default_rng(11)
    _say("spilled dimensions")
    D0, D1 = 3_000_000, 1_500_000
    dvs = [None] * 4
    dvs[0] = po.build_dimvec(rng.integers(0, 100, D0, dtype=np.int32),
                             rng.integers(0, 100, D0, dtype=np.int32), 0, 50, 1, 1)
    dvs[2] = po.build_dimvec(rng.integers(0, 100, D1, dtype=np.int32),
                             rng.integers(0, 3, D1, dtype=np.int32), 0, 30, 1, 1)
    rows = 2_000_000
    fks = [rng.integers(0, D0, rows, dtype=np.int32), None,
           rng.integers(0, D1, rows, dtype=np.int32), None]
    p = po.plan([2, 0], [100, 1], 300)
    m = rng.integers(1, 1_000_000, rows, dtype=np.int32)
    check_all(ctx, p, dvs, fks, m, None, ncpu)


def test_column_wise_split_entry_points(ctx, ncpu):
    import torch

    rng = np.random.default_rng(21)
    _say("join/agg split entry points")
    plan_np, dv, fk, m1, m2 = random_star(rng, 3, 123_457)
    star = po.Star(dv, fk, m1, m2)
    es, ec, _ = po.run("rowwise", plan_np, star)
    egid = po.join_cwm_sv(plan_np, star)
    d_dv = [_dev(d, torch) for d in dv]
    d_fk = [_dev(f, torch) for f in fk]
    plan = oc.plan_from(list(plan_np.orders)[: plan_np.dimvec_nums],
                        list(plan_np.factor)[: plan_np.dimvec_nums], int(plan_np.group_nums))
    cols = oc.Context.columns(d_dv, d_fk, _dev(m1, torch), _dev(m2, torch), rows=m1.size)
    q = ctx.query(plan, cols)
    G = int(plan.group_nums)
    gvec = torch.empty(m1.size, dtype=torch.int32, device="cuda")
    q.join_cwm_sv(gvec)
    osum = torch.empty(G, dtype=torch.int64, device="cuda")
    ocnt = torch.empty(G, dtype=torch.int64, device="cuda")
    q.agg_cwm_sv(gvec, osum, ocnt)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(gvec.cpu().numpy(), egid)
    np.testing.assert_array_equal(osum.cpu().numpy(), es)
    np.testing.assert_array_equal(ocnt.cpu().numpy(), ec)
    q.join_cwm_dv()
    q.agg_cwm_dv(osum, ocnt)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(osum.cpu().numpy(), es)
    np.testing.assert_array_equal(ocnt.cpu().numpy(), ec)
    q.close()


def test_one_shot_reference_named(ctx, ncpu):
    import torch

    rng = np.random.default_rng(31)
    _say("one-shot reference-named entry points")
    plan_np, dv, fk, m1, m2 = random_star(rng, 4, 77_001, two_measures=True)
    star = po.Star(dv, fk, m1, m2)
    es, ec, _ = po.run("vwm_dv", plan_np, star, nthreads=ncpu)
    plan = oc.plan_from(list(plan_np.orders)[: plan_np.dimvec_nums],
                        list(plan_np.factor)[: plan_np.dimvec_nums], int(plan_np.group_nums))
    cols = oc.Context.columns([_dev(d, torch) for d in dv], [_dev(f, torch) for f in fk],
                              _dev(m1, torch), _dev(m2, torch), rows=m1.size)
    G = int(plan.group_nums)
    for s in STRATS:
        osum = torch.empty(G, dtype=torch.int64, device="cuda")
        ocnt = torch.empty(G, dtype=torch.int64, device="cuda")
        total = ctx.run(s, plan, cols, osum, ocnt)
        assert total == int(ec.sum())
        np.testing.assert_array_equal(osum.cpu().numpy(), es)
        np.testing.assert_array_equal(ocnt.cpu().numpy(), ec)
    assert oc.legacy_count(es) == po.legacy_count(es)


def test_invalid_inputs_rejected(ctx):
    import torch

    _say("invalid inputs")
    dv = torch.tensor([0, 1, -5, 127], dtype=torch.int8, device="cuda")
    fk = torch.zeros(16, dtype=torch.int32, device="cuda")
    m = torch.ones(16, dtype=torch.int32, device="cuda")
    cols = oc.Context.columns([dv, None, None, None], [fk, None, None, None], m)
    with pytest.raises(oc.OlapError) as e:
        ctx.query(oc.plan_from([0], [1], 4), cols)
    assert e.value.status == oc.E_INVALID
    dv2 = torch.tensor([0, 1, 2, 127], dtype=torch.int8, device="cuda")
    cols = oc.Context.columns([dv2, None, None, None], [fk, None, None, None], m)
    with pytest.raises(oc.OlapError):
        ctx.query(oc.plan_from([0], [1], 2), cols)  # code 2 reaches slot 2 of 2
    with pytest.raises(oc.OlapError):
        ctx.query(oc.plan_from([0, 0], [1, 1], 4), cols)  # repeated table


def test_c1_full_size(ctx, ncpu):
    """The 1-dimension star at full size: 6M fact rows, one 100,000-row dimension."""
    import torch

    cfg = wl.STAR_1DIM_6M
    _say(f"{cfg.name}: generating on device")
    ds = wl.DeviceStar(ctx, cfg)
    dimvecs, _, _, fks, m1, m2 = cpu_star(cfg)
    for d in cfg.dims:
        np.testing.assert_array_equal(ds.dimvecs[d.slot].cpu().numpy(), dimvecs[d.slot])
    p = po.make_plan(cfg.groups_by_slot(), cfg.bitmap_by_slot(), cfg.sele_by_slot())
    es, ec, _ = po.run("rowwise", p, po.Star(dimvecs, fks, m1, m2), nthreads=ncpu)
    G = int(ds.plan.group_nums)
    q = ctx.query(ds.plan, ds.cols)
    for s in STRATS:
        _say(f"  {cfg.name} {s}")
        osum = torch.empty(G, dtype=torch.int64, device="cuda")
        ocnt = torch.empty(G, dtype=torch.int64, device="cuda")
        q.run(s, osum, ocnt)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(ocnt.cpu().numpy(), ec)
        np.testing.assert_array_equal(osum.cpu().numpy(), es)
    q.close()


def test_c2_full_size(ctx, ncpu):
    """The 4-dimension star at full size: 60M fact rows, four dimensions, G=875."""
    import torch

    cfg = wl.STAR_4DIM_60M
    _say(f"{cfg.name}: generating on device")
    ds = wl.DeviceStar(ctx, cfg)
    dimvecs, _, _, fks, m1, m2 = cpu_star(cfg)
    p = po.make_plan(cfg.groups_by_slot(), cfg.bitmap_by_slot(), cfg.sele_by_slot())
    assert list(p.orders) == list(ds.plan.orders) and list(p.factor) == list(ds.plan.factor)
    es, ec, _ = po.run("vwm_dv", p, po.Star(dimvecs, fks, m1, m2), nthreads=ncpu)
    assert ec.sum() > 0
    G = int(ds.plan.group_nums)
    q = ctx.query(ds.plan, ds.cols)
    for s in STRATS:
        _say(f"  {cfg.name} {s}")
        osum = torch.empty(G, dtype=torch.int64, device="cuda")
        ocnt = torch.empty(G, dtype=torch.int64, device="cuda")
        q.run(s, osum, ocnt)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(ocnt.cpu().numpy(), ec)
        np.testing.assert_array_equal(osum.cpu().numpy(), es)
    q.close()


def test_starjoin_substep(ctx):
    import torch

    rng = np.random.default_rng(41)
    _say("starjoin sub-step")
    rows = 300_001
    dv, fk = [], []
    for D in (3000, 200, 20000, 2555):
        d = rng.integers(0, 16, D).astype(np.int8)
        d[rng.random(D) < 0.4] = 127
        dv.append(d)
        fk.append(rng.integers(0, D + 1, rows, dtype=np.int32))
    cnt, res = po.starjoin_count(dv, fk, want_result=True)
    ddv = [_dev(d, torch) for d in dv]
    dfk = [_dev(f, torch) for f in fk]
    for s in STRATS:
        r = torch.empty(rows, dtype=torch.int32, device="cuda") if s == "cwm_sv" else None
        got = ctx.starjoin(s, ddv, dfk, result=r)
        assert got == cnt, s
        if r is not None:
            np.testing.assert_array_equal(r.cpu().numpy(), res)


@pytest.mark.parametrize("G", [32, 1024, 16384, 1 << 18])
def test_vector_group_substep(ctx, G):
    import torch

    rng = np.random.default_rng(G)
    _say(f"vector_group G={G}")
    n = 2_000_003
    key = rng.integers(0, G + 2, n, dtype=np.int32)  # includes 0 and G+1: bad keys
    m1 = rng.integers(-1000, 1000, n, dtype=np.int32)
    m2 = rng.integers(0, 5, n, dtype=np.int32)
    tot, gv, bad = po.vector_group(key, m1, m2, G)
    out = torch.empty(G, dtype=torch.int64, device="cuda")
    gt, gbad = ctx.vector_group(_dev(key, torch), _dev(m1, torch), _dev(m2, torch), out)
    assert gt == tot and gbad == bad
    np.testing.assert_array_equal(out.cpu().numpy(), gv)


def test_agg_expr_substep(ctx):
    import torch

    rng = np.random.default_rng(3)
    n = 1 << 20
    p = rng.integers(5, 21, n, dtype=np.int32)
    q = rng.integers(1, 11, n, dtype=np.int32)
    t = rng.integers(1, 21, n, dtype=np.int32)
    assert ctx.agg_expr(_dev(p, torch), _dev(q, torch), _dev(t, torch)) == po.agg_expr(p, q, t)


def test_every_row_shape_matches_oracle(ctx, ncpu):
    """Each launch shape of the row-wise scan, and the autotuner's pick."""
    import torch

    rng = np.random.default_rng(51)
    plan_np, dv, fk, m1, m2 = random_star(rng, 4, 1_234_567, oob=True)
    star = po.Star(dv, fk, m1, m2)
    es, ec, _ = po.run("rowwise", plan_np, star, nthreads=ncpu)
    plan = oc.plan_from(list(plan_np.orders)[: plan_np.dimvec_nums],
                        list(plan_np.factor)[: plan_np.dimvec_nums], int(plan_np.group_nums))
    cols = oc.Context.columns([_dev(d, torch) for d in dv], [_dev(f, torch) for f in fk],
                              _dev(m1, torch), _dev(m2, torch), rows=m1.size)
    G = int(plan.group_nums)
    q = ctx.query(plan, cols)
    osum = torch.empty(G, dtype=torch.int64, device="cuda")
    ocnt = torch.empty(G, dtype=torch.int64, device="cuda")
    for shape in range(ctx.lib.olap_num_shapes(oc.ROWWISE)):
        _say(f"  row shape {shape}")
        q.set_shape(shape)
        q.run("rowwise", osum, ocnt)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(ocnt.cpu().numpy(), ec, err_msg=f"shape {shape}")
        np.testing.assert_array_equal(osum.cpu().numpy(), es, err_msg=f"shape {shape}")
    shape, threads, rpt, ms = q.autotune("rowwise", osum, ocnt)
    _say(f"  autotune picked shape {shape} ({threads} threads x {rpt} rows) {ms:.3f} ms")
    np.testing.assert_array_equal(ocnt.cpu().numpy(), ec)
    np.testing.assert_array_equal(osum.cpu().numpy(), es)
    q.close()


PIPE_CASES = {
    # first step passes 1% (its later keys are not streamed), two measures
    "sparse_first": dict(rows=1_234_567, thr=[1, 40, 60, 30], groups=[3, 4, 5, 2], m2=True),
    # every step passes about half (step 1's keys streamed with step 0's)
    "dense": dict(rows=1_000_003, thr=[50, 55, 45, 60], groups=[7, 3, 5, 4], m2=False),
    # fewer rows than one tile, then a few tiles and a ragged end
    "tiny": dict(rows=5, thr=[70, 80, 90, 60], groups=[2, 3, 2, 2], m2=True),
    "few_tiles": dict(rows=3 * 8192 + 3, thr=[30, 80, 90, 60], groups=[2, 3, 2, 2], m2=False),
    # no measure column: counts only
    "count_only": dict(rows=777_777, thr=[20, 50, 50, 50], groups=[3, 3, 3, 3], m2=None),
    # group table beyond LDS: global atomics, then behind the hot-group cache
    "wide_groups": dict(rows=900_001, thr=[60, 70, 80, 90], groups=[200, 150, 30, 4], m2=False),
}


@pytest.mark.parametrize("ndims", [1, 2, 3, 4])
@pytest.mark.parametrize("case", list(PIPE_CASES))
def test_row_shapes_every_pipeline_depth(ctx, ncpu, ndims, case):
    """Every row-wise launch shape (the pipelined ones issue each dependent
    fetch a tile ahead, with a depth of ndims - 1 or ndims stages) against
    the oracle: streamed or unstreamed second step, one or two measures,
    counts only, ragged ends, group tables in LDS or in HBM."""
    import torch

    c = PIPE_CASES[case]
    rng = np.random.default_rng(700 + 10 * ndims + len(case))
    rows = c["rows"]
    slots = list(rng.permutation(4)[:ndims])
    groups, bitmap, sele = [0] * 4, [0] * 4, [0.0] * 4
    dv, fk = [None] * 4, [None] * 4
    for i, t in enumerate(slots):
        D = int(rng.integers(1000, 40_000))
        g = c["groups"][i]
        groups[t], bitmap[t], sele[t] = g, 1, c["thr"][i] / 100.0
        f = rng.integers(0, 100, D, dtype=np.int32)
        dv[t] = po.build_dimvec(f, rng.integers(0, g, D, dtype=np.int32), -(2 ** 31), c["thr"][i], 1,
                                1 if g < 127 else 4)
        fk[t] = rng.integers(-2, D + D // 16, rows, dtype=np.int32)
    plan_np = po.make_plan(groups, bitmap, sele)
    count_only = c["m2"] is None
    m1 = None if count_only else rng.integers(-2 ** 31, 2 ** 31 - 1, rows, dtype=np.int32)
    m2 = rng.integers(-2 ** 31, 2 ** 31 - 1, rows, dtype=np.int32) if c["m2"] else None
    ref_m1 = np.zeros(rows, np.int32) if count_only else m1
    es, ec, _ = po.run("rowwise", plan_np, po.Star(dv, fk, ref_m1, m2), nthreads=ncpu)
    plan = oc.plan_from(list(plan_np.orders)[: plan_np.dimvec_nums],
                        list(plan_np.factor)[: plan_np.dimvec_nums], int(plan_np.group_nums))
    cols = oc.Context.columns([None if d is None else _dev(d, torch) for d in dv],
                              [None if f is None else _dev(f, torch) for f in fk],
                              None if m1 is None else _dev(m1, torch),
                              None if m2 is None else _dev(m2, torch), rows=rows)
    G = int(plan.group_nums)
    q = ctx.query(plan, cols)
    osum = torch.empty(G, dtype=torch.int64, device="cuda")
    ocnt = torch.empty(G, dtype=torch.int64, device="cuda")
    layouts = [0, 2] if case == "wide_groups" else [-1]
    for layout in layouts:
        q.set_aggregation(layout)
        for shape in range(ctx.lib.olap_num_shapes(oc.ROWWISE)):
            _say(f"  {case} ndims={ndims} G={G} layout={layout} shape {shape}")
            q.set_shape(shape)
            q.run("rowwise", None if count_only else osum, ocnt)
            torch.cuda.synchronize()
            pipe = ctx.lib.olap_shape_pipeline(oc.ROWWISE, shape)
            ran = q.pipeline()
            if pipe in (3, 4, 5, 6):  # queues that do not fit the LDS the query leaves take the pipelined form
                assert ran["survivor_queue"] or (layout == 2 and ran["fetch_ahead"])
            else:
                assert ran["fetch_ahead"] == (pipe in (1, 2)) and not ran["survivor_queue"]
            if pipe in (2, 5, 6):
                assert not ran["second_step_streamed"]
            np.testing.assert_array_equal(ocnt.cpu().numpy(), ec, err_msg=f"shape {shape} counts")
            if not count_only:
                np.testing.assert_array_equal(osum.cpu().numpy(), es, err_msg=f"shape {shape} sums")
    assert int(ec.sum()) > 0 or rows < 100
    q.close()
