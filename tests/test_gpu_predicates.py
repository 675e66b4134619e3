"""General dimension predicates on the GPU (olap_build_dimvec_program and
the batch mask builder), bit-exact against the oracle (whose evaluation is
pinned by tests/test_predicates.py): random range / IN / hop programs with
AND / OR and negation, a two-hop snowflake, and batches whose steps carry
programs against the same queries run as plain star joins and against the
oracle's star join."""
import numpy as np
import pytest

from memory_database_amd import olapcore as oc
from memory_database_amd import workloads as wl
from oracle import pyoracle as po
from tests.predicate_cases import random_program

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def device_terms(terms):
    out = []
    for t in terms:
        t = list(t)
        words = []
        while t[0] in ("and", "or", "not"):
            words.append(t.pop(0))
        if t[0] == "in":   # IN values stay on the host
            out.append(tuple(words + ["in", _dev(t[1]), t[2]]))
        elif t[0] == "range":
            out.append(tuple(words + ["range", _dev(t[1]), t[2], t[3]]))
        else:
            out.append(tuple(words + ["hop", _dev(t[1]), _dev(t[2]), t[3]]))
    return out


@pytest.mark.parametrize("seed", range(16))
def test_random_programs_match_oracle(ctx, seed):
    import torch

    rng = np.random.default_rng(900 + seed)
    rows = int(rng.integers(1, 300_000))
    pdt = np.int8 if seed % 2 else np.int32
    prow = int(rng.integers(1, 100_000))
    parent = rng.integers(0, 100, prow).astype(pdt)
    parent[rng.random(prow) < 0.3] = 127 if pdt == np.int8 else -1
    terms = random_program(rng, rows, parent, nterms=int(rng.integers(1, 8)))
    if seed % 4 == 0:  # a long IN list, the eighth term at most
        col = rng.integers(0, 20_000, rows, dtype=np.int32)
        terms.append(("or", "in", col, rng.integers(0, 20_000, 4096, dtype=np.int32)))
    words = [[w for w in t if isinstance(w, str)] for t in terms]
    hops = [i for i, w in enumerate(words) if "hop" in w and "not" not in w]
    kind = ["none", "column", "hop"][seed % 3]
    if kind == "hop" and not hops:
        kind = "column"
    group = {"none": oc.GROUP_NONE, "column": oc.GROUP_COLUMN, "hop": oc.GROUP_HOP}[kind]
    gcol = rng.integers(0, 5000, rows, dtype=np.int32)
    gmod = 120 if seed % 2 else int(rng.integers(0, 3000))
    gterm = hops[0] if hops else 0
    code_bytes = 1 if (seed % 2 and kind != "hop") or (kind == "hop" and pdt == np.int8) else 4
    _say(f"program seed={seed} rows={rows} terms={len(terms)} group={kind} bytes={code_bytes}")
    want, bad = po.build_dimvec_program(terms, rows, code_bytes, group=group, group_col=gcol,
                                        group_mod=gmod, group_term=gterm)
    prog = oc.DimProgram.of(device_terms(terms), group=group, group_col=_dev(gcol), group_mod=gmod,
                            group_term=gterm)
    out = torch.empty(rows, dtype=torch.int8 if code_bytes == 1 else torch.int32, device="cuda")
    if bad:
        with pytest.raises(oc.OlapError):
            ctx.build_dimvec_program(prog, rows, out)
        return
    ctx.build_dimvec_program(prog, rows, out)
    np.testing.assert_array_equal(out.cpu().numpy(), want)


def test_int8_codes_out_of_range_are_refused(ctx):
    import torch

    f = _dev(np.arange(1000, dtype=np.int32))
    prog = oc.DimProgram.of([("range", f, 0, 1000)], group=oc.GROUP_COLUMN, group_col=f, group_mod=0)
    out = torch.empty(1000, dtype=torch.int8, device="cuda")
    with pytest.raises(oc.OlapError) as e:
        ctx.build_dimvec_program(prog, 1000, out)
    assert e.value.status == oc.E_INVALID


def test_two_hop_snowflake_matches_oracle(ctx):
    import torch

    from tests.test_predicates import q5_schema

    rng = np.random.default_rng(55)
    r_code, n_regionkey, s_nationkey, c_nationkey, o_custkey, o_orderdate = q5_schema(rng)
    target = int(r_code[1])
    region, _ = po.build_dimvec_program([("in", r_code, [target])], 5, 4)
    nation, _ = po.build_dimvec_program([("hop", n_regionkey, region, 0)], 25, 4, group=1,
                                        group_col=np.arange(25, dtype=np.int32))
    cust, _ = po.build_dimvec_program([("hop", c_nationkey, nation, 0)], c_nationkey.size, 4,
                                      group=2, group_term=0)
    orders, _ = po.build_dimvec_program([("range", o_orderdate, 19940101, 19950101),
                                         ("and", "hop", o_custkey, cust, -1)], o_custkey.size, 4,
                                        group=2, group_term=1)
    d_region = torch.empty(5, dtype=torch.int32, device="cuda")
    ctx.build_dimvec_program(oc.DimProgram.of([("in", _dev(r_code), [target])]), 5, d_region)
    d_nation = torch.empty(25, dtype=torch.int32, device="cuda")
    ctx.build_dimvec_program(oc.DimProgram.of([("hop", _dev(n_regionkey), d_region, 0)],
                                              group=oc.GROUP_COLUMN,
                                              group_col=_dev(np.arange(25, dtype=np.int32))),
                             25, d_nation)
    d_cust = torch.empty(c_nationkey.size, dtype=torch.int32, device="cuda")
    ctx.build_dimvec_program(oc.DimProgram.of([("hop", _dev(c_nationkey), d_nation, 0)],
                                              group=oc.GROUP_HOP, group_term=0),
                             c_nationkey.size, d_cust)
    d_orders = torch.empty(o_custkey.size, dtype=torch.int32, device="cuda")
    ctx.build_dimvec_program(oc.DimProgram.of([("range", _dev(o_orderdate), 19940101, 19950101),
                                               ("and", "hop", _dev(o_custkey), d_cust, -1)],
                                              group=oc.GROUP_HOP, group_term=1),
                             o_custkey.size, d_orders)
    for name, d, h in (("region", d_region, region), ("nation", d_nation, nation),
                       ("customer", d_cust, cust), ("orders", d_orders, orders)):
        np.testing.assert_array_equal(d.cpu().numpy(), h, err_msg=name)


def test_batch_with_programs_equals_plain_and_oracle(ctx, ncpu):
    """A batch of 24 queries over a 2M-row star where some steps carry a
    program (IN lists, OR of ranges, a NOT); every query equals the plain
    star join over vectors built by the program ANDed with the step's range
    and grouped by attr % g, and the oracle's star join."""
    import torch

    from tests.test_gpu_batch import host_star

    cfg = wl.scaled(wl.STAR_4DIM_600M, 2_000_003, name="pred_batch")
    bs = wl.DeviceBatchStar(ctx, cfg)
    h_filters, h_attrs, h_fks, h_m1 = host_star(cfg, cfg.rows)
    rng = np.random.default_rng(77)
    tmpl = wl.batch_queries(cfg, 24, seed=19, max_groups=5000)
    queries, host_progs = [], []
    for qi, (steps, factor, G) in enumerate(tmpl):
        progs, hp = [], []
        for (t, lo, hi, g) in steps:
            if rng.integers(0, 2):
                fcol = h_filters[t]
                acol = h_attrs[t]
                terms = [("in", fcol, rng.integers(0, 100, int(rng.integers(1, 30)), dtype=np.int32)),
                         ("or", "range", acol, 0, int(rng.integers(1, 1 << 19))),
                         ("and", "not", "range", fcol, 40, 45)]
                hp.append(terms)
                progs.append(oc.DimProgram.of(device_terms(terms)))
            else:
                hp.append(None)
                progs.append(None)
        host_progs.append(hp)
        queries.append(oc.BatchQuery.make(steps, factor, G, programs=progs))
    b = ctx.batch(bs.star, queries)
    osum = torch.empty(b.total_slots, dtype=torch.int64, device="cuda")
    ocnt = torch.empty_like(osum)
    b.run(osum, ocnt)
    torch.cuda.synchronize()
    bsum, bcnt = osum.cpu().numpy(), ocnt.cpu().numpy()
    for qi, (steps, factor, G) in enumerate(tmpl):
        dvs_h = [None] * 4
        dvs_d = [None] * 4
        for si, (t, lo, hi, g) in enumerate(steps):
            base = host_progs[qi][si] or []
            terms = list(base) + [("and", "range", h_filters[t], lo, hi)]
            grp = dict(group=1, group_col=h_attrs[t], group_mod=g) if g > 0 else dict(group=0)
            dvs_h[t], _ = po.build_dimvec_program(terms, h_filters[t].size, 4, **grp)
            dvs_d[t] = _dev(dvs_h[t])
        plan = oc.plan_from([s[0] for s in steps], factor, G)
        cols = oc.Context.columns(dvs_d, bs.fks, bs.m1, None, rows=cfg.rows)
        q = ctx.query(plan, cols)
        ps = torch.empty(G, dtype=torch.int64, device="cuda")
        pc = torch.empty_like(ps)
        q.run("rowwise", ps, pc)
        torch.cuda.synchronize()
        q.close()
        lo_, hi_ = int(b.offsets[qi]), int(b.offsets[qi + 1])
        np.testing.assert_array_equal(bcnt[lo_:hi_], pc.cpu().numpy(), err_msg=f"query {qi} counts")
        np.testing.assert_array_equal(bsum[lo_:hi_], ps.cpu().numpy(), err_msg=f"query {qi} sums")
        if qi % 4 == 0:
            es, ec, _ = po.run("rowwise", po.plan([s[0] for s in steps], factor, G),
                               po.Star(dvs_h, h_fks, h_m1, None), nthreads=ncpu)
            np.testing.assert_array_equal(bcnt[lo_:hi_], ec, err_msg=f"query {qi} vs oracle")
            np.testing.assert_array_equal(bsum[lo_:hi_], es, err_msg=f"query {qi} vs oracle")
    b.close()
