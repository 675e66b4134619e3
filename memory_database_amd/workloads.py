"""Synthetic star schemas of the benchmark configurations, generated on the
device through libolapcore's counter-based generator.

Every column is a pure function of (seed, stream id, row), so any slice of
any table can be regenerated on the host for checking. Stream ids: dimension
slot t uses 100+t (filter) and 200+t (group); the fact table uses 300+t for
the foreign key into slot t and 400/401 for the measures.

Table slots follow the reference's main(): 0 customer, 1 supplier, 2 part,
3 date (code/OLAPcore/test_multi_compute_operator.cpp:1930-1953).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

DEFAULT_SEED = 20251203


@dataclass
class Dim:
    slot: int            # table slot 0..3
    rows: int            # dimension rows D
    pred_hi: int         # filter U[0,100) passes when < pred_hi (selectivity pred_hi/100)
    groups: int          # group column U[0, groups)
    bitmap: int = 1      # 1: group-bearing, 0: filter-only
    pred_lo: int = -(2 ** 31)

    @property
    def sele(self) -> float:
        return self.pred_hi / 100.0


@dataclass
class StarConfig:
    name: str
    rows: int
    dims: List[Dim]
    m_lo: int            # measure U[m_lo, m_hi]
    m_hi: int
    zipf_theta: float = 0.0
    seed: int = DEFAULT_SEED
    two_measures: bool = False
    notes: str = ""

    def groups_by_slot(self):
        g = [0, 0, 0, 0]
        for d in self.dims:
            g[d.slot] = d.groups
        return g

    def bitmap_by_slot(self):
        b = [0, 0, 0, 0]
        for d in self.dims:
            b[d.slot] = d.bitmap
        return b

    def sele_by_slot(self):
        s = [0.0, 0.0, 0.0, 0.0]
        for d in self.dims:
            s[d.slot] = d.sele
        return s

    def fact_bytes_per_row(self) -> int:
        return 4 * (len(self.dims) + (2 if self.two_measures else 1))


# Plain 1-dim star-join group-by on the reference CPU path: 6M fact rows, FK
# U[0,100000), measure U[1,100]; 100,000-row dimension, filter < 10, 16 groups.
STAR_1DIM_6M = StarConfig("star_1dim_6M", 6_000_000, [Dim(0, 100_000, 10, 16)], 1, 100)

# Plain 4-dim star-join group-by: 60M fact rows; dimensions date 2,556 /
# customer 30,000 / part 200,000 / supplier 2,000 rows with selectivities
# 0.5 / 0.2 / 0.04 / 0.2 and group cardinalities 7 / 25 / 1 / 5 (G = 875);
# measure U[1000, 1000000].
STAR_DIMS = [Dim(3, 2_556, 50, 7), Dim(0, 30_000, 20, 25), Dim(2, 200_000, 4, 1),
           Dim(1, 2_000, 20, 5)]
STAR_4DIM_60M = StarConfig("star_4dim_60M", 60_000_000, STAR_DIMS, 1000, 1_000_000)

# The same star at 600M fact rows (12 GB of fact columns).
STAR_4DIM_600M = StarConfig("star_4dim_600M", 600_000_000, STAR_DIMS, 1000, 1_000_000)

CONFIGS = {c.name: c for c in (STAR_1DIM_6M, STAR_4DIM_60M, STAR_4DIM_600M)}

# Sweep of the 4-dimension star: Zipf skew of every foreign key, total group
# count G (split over date / customer / part / supplier) and the selectivity
# of every dimension predicate.
SWEEP_THETA = (0.0, 0.5, 1.0)
SWEEP_GROUPS = {100: (4, 25, 1, 1), 10_000: (10, 100, 1, 10), 1_000_000: (10, 100, 10, 100)}
SWEEP_SELE = (1, 10, 50)   # percent, every dimension


def sweep_configs(rows: int = 600_000_000, thetas=SWEEP_THETA, groups=tuple(SWEEP_GROUPS),
                  seles=SWEEP_SELE):
    """Star configurations of the sweep, grouped so that configurations with
    the same theta are adjacent (they share the fact columns)."""
    out = []
    for th in thetas:
        for G in groups:
            gd, gc, gp, gs = SWEEP_GROUPS[G]
            for sel in seles:
                dims = [Dim(3, 2_556, sel, gd), Dim(0, 30_000, sel, gc), Dim(2, 200_000, sel, gp),
                        Dim(1, 2_000, sel, gs)]
                name = f"sweep_4dim_{rows // 1_000_000}M_theta{th:g}_G{G}_sel{sel}"
                out.append(StarConfig(name, rows, dims, 1000, 1_000_000, zipf_theta=th))
    return out


def config_by_name(name: str) -> StarConfig:
    """A named configuration: the three plain workloads or a point of the
    600M-row sweep (e.g. sweep_4dim_600M_theta0_G1000000_sel50)."""
    if name in CONFIGS:
        return CONFIGS[name]
    for c in sweep_configs():
        if c.name == name:
            return c
    raise KeyError(f"unknown configuration {name}")


def scaled(cfg: StarConfig, rows: int, name: Optional[str] = None, **kw) -> StarConfig:
    d = dict(cfg.__dict__)
    d.update(kw)
    d["rows"] = rows
    d["name"] = name or f"{cfg.name}@{rows}"
    return StarConfig(**d)


class DeviceStar:
    """A StarConfig materialised in HBM (torch tensors on the ROCm device)."""

    def __init__(self, ctx, cfg: StarConfig, device: str = "cuda", stream=None,
                 code_bytes: int = 1, fact: Optional["DeviceStar"] = None):
        """fact: a DeviceStar of the same rows, seed, dimension sizes and skew
        whose foreign keys and measures are reused instead of regenerated."""
        import torch

        from . import olapcore as oc

        self.cfg = cfg
        self.ctx = ctx
        i32 = torch.int32
        self.filters = [None] * 4
        self.groups = [None] * 4
        self.dimvecs = [None] * 4
        self.fks = [None] * 4
        dt = torch.int8 if code_bytes == 1 else torch.int32
        zipf_thr = {}
        for d in cfg.dims:
            f = torch.empty(d.rows, dtype=i32, device=device)
            g = torch.empty(d.rows, dtype=i32, device=device)
            ctx.gen_uniform(f, cfg.seed, 100 + d.slot, 0, 100, stream=stream)
            ctx.gen_uniform(g, cfg.seed, 200 + d.slot, 0, d.groups, stream=stream)
            dv = torch.empty(d.rows, dtype=dt, device=device)
            ctx.build_dimvec(f, g, d.pred_lo, d.pred_hi, d.bitmap, dv,
                             group_mod=d.groups if code_bytes == 1 else 0, stream=stream)
            self.filters[d.slot], self.groups[d.slot], self.dimvecs[d.slot] = f, g, dv
            if fact is not None:
                self.fks[d.slot] = fact.fks[d.slot]
                continue
            fk = torch.empty(cfg.rows, dtype=i32, device=device)
            if cfg.zipf_theta > 0.0:
                thr = torch.from_numpy(oc.zipf_table(d.rows, cfg.zipf_theta).view("int64")).to(device)
                zipf_thr[d.slot] = thr
                ctx.gen_zipf(fk, cfg.seed, 300 + d.slot, thr, stream=stream)
            else:
                ctx.gen_uniform(fk, cfg.seed, 300 + d.slot, 0, d.rows, stream=stream)
            self.fks[d.slot] = fk
        if fact is not None:
            self.m1 = fact.m1
        else:
            self.m1 = torch.empty(cfg.rows, dtype=i32, device=device)
            ctx.gen_uniform(self.m1, cfg.seed, 400, cfg.m_lo, cfg.m_hi + 1, stream=stream)
        self.m2 = None
        if cfg.two_measures:
            self.m2 = torch.empty(cfg.rows, dtype=i32, device=device)
            ctx.gen_uniform(self.m2, cfg.seed, 401, cfg.m_lo, cfg.m_hi + 1, stream=stream)
        self._zipf = zipf_thr  # generator tables stay alive with the columns
        self.plan = oc.make_plan(cfg.groups_by_slot(), cfg.bitmap_by_slot(), cfg.sele_by_slot())
        self.cols = oc.Context.columns(self.dimvecs, self.fks, self.m1, self.m2, rows=cfg.rows)

    def fact_bytes(self) -> int:
        return self.cfg.rows * self.cfg.fact_bytes_per_row()


# ---------------------------------------------------------------- batched form

ATTR_HI = 1 << 20   # group attribute of every dimension: U[0, 2^20), stream 500+t


def batch_queries(cfg: StarConfig, nq: int = 64, seed: int = 7, max_groups: int = 100_000):
    """Seeded mixed star-join group-by queries over cfg's dimensions: each
    draws 1-4 distinct dimensions, per dimension a predicate range lo <= filter
    < hi on the U[0,100) filter column giving 1%-50% selectivity, and a total
    group cardinality G uniform over [1, max_groups] split over its dimensions
    as GROUP BY attr % g_i (G = prod g_i). Steps are ordered most selective
    first. Returns [(steps [(table, lo, hi, g)], factor, G)]."""
    import numpy as np

    rng = np.random.default_rng(seed)
    slots = [d.slot for d in cfg.dims]
    out = []
    for _ in range(nq):
        k = int(rng.integers(1, min(4, len(slots)) + 1))
        tables = [int(t) for t in rng.permutation(slots)[:k]]
        widths = [int(rng.integers(1, 51)) for _ in range(k)]
        target = int(rng.integers(1, max_groups + 1))
        g = []
        rest = target
        for i in range(k):
            if i == k - 1:
                gi = max(1, rest)
            else:
                gi = max(1, int(round(target ** (1.0 / k))))
                rest = max(1, rest // gi)
            g.append(gi)
        steps = []
        for t, w, gi in zip(tables, widths, g):
            lo = int(rng.integers(0, 101 - w))
            steps.append((t, lo, lo + w, gi))
        steps.sort(key=lambda s: s[2] - s[1])
        gs = [s[3] for s in steps]
        factor = [int(np.prod(gs[i + 1:], dtype=np.int64)) for i in range(k)]
        G = int(np.prod(gs, dtype=np.int64))
        out.append((steps, factor, G))
    return out


class DeviceBatchStar:
    """The shared star of the batched form in HBM: cfg's fact table plus per
    dimension a filter column U[0,100) and a group attribute U[0, 2^20)."""

    def __init__(self, ctx, cfg: StarConfig, device: str = "cuda", stream=None, fact=None):
        import torch

        from . import olapcore as oc

        self.cfg = cfg
        self.ctx = ctx
        i32 = torch.int32
        self.filters = [None] * 4
        self.attrs = [None] * 4
        self.fks = [None] * 4
        for d in cfg.dims:
            f = torch.empty(d.rows, dtype=i32, device=device)
            a = torch.empty(d.rows, dtype=i32, device=device)
            ctx.gen_uniform(f, cfg.seed, 100 + d.slot, 0, 100, stream=stream)
            ctx.gen_uniform(a, cfg.seed, 500 + d.slot, 0, ATTR_HI, stream=stream)
            self.filters[d.slot], self.attrs[d.slot] = f, a
            if fact is not None:
                self.fks[d.slot] = fact.fks[d.slot]
            else:
                fk = torch.empty(cfg.rows, dtype=i32, device=device)
                ctx.gen_uniform(fk, cfg.seed, 300 + d.slot, 0, d.rows, stream=stream)
                self.fks[d.slot] = fk
        if fact is not None:
            self.m1 = fact.m1
        else:
            self.m1 = torch.empty(cfg.rows, dtype=i32, device=device)
            ctx.gen_uniform(self.m1, cfg.seed, 400, cfg.m_lo, cfg.m_hi + 1, stream=stream)
        self.star = oc.Context.star(self.filters, self.attrs, self.fks, self.m1, None, rows=cfg.rows)

    def plain_query(self, steps, factor, G, stream=None):
        """The same query as a plain star join: int32 dimension vectors built
        with olap_build_dimvec, and its plan."""
        import torch

        from . import olapcore as oc

        dvs = [None] * 4
        for (t, lo, hi, g) in steps:
            dv = torch.empty(self.filters[t].numel(), dtype=torch.int32, device=self.filters[t].device)
            self.ctx.build_dimvec(self.filters[t], self.attrs[t], lo, hi, 1 if g > 0 else 0, dv,
                                  group_mod=g, stream=stream)
            dvs[t] = dv
        plan = oc.plan_from([s[0] for s in steps], factor, G)
        cols = oc.Context.columns(dvs, self.fks, self.m1, None, rows=self.cfg.rows)
        return plan, cols


# ---------------------------------------------------------------- TPC-H Q5

REGIONS = ["AFRICA", "AMERICA", "ASIA", "EUROPE", "MIDDLE EAST"]
# n_regionkey of the 25 TPC-H nations in n_nationkey order (the TPC-H
# specification's NATION table)
NATION_REGION = [0, 1, 1, 1, 4, 0, 3, 3, 2, 2, 4, 4, 2, 4, 0, 0, 0, 1, 2, 3, 4, 2, 3, 3, 1]
LINES_PER_ORDER = 4


@dataclass
class Q5Config:
    """A TPC-H-shaped snowflake for Q5 at scale factor sf: 10,000 suppliers,
    150,000 customers and 1,500,000 orders per unit of sf, 4 lineitems per
    order (6M per unit of sf), lineitems in order-key order as TPC-H lays
    them out. Columns are uniform draws of the counter-based generator:
    s_nationkey, c_nationkey U[0,25); o_custkey U[1,C]; o_orderdate
    yyyymmdd over 1992-1998 (days 1-28); l_suppkey U[1,S]; l_extendedprice
    in cents U[90000, 10494950]; l_discount in percent U[0,10]."""
    sf: float
    seed: int = DEFAULT_SEED
    name: str = ""

    @property
    def suppliers(self) -> int:
        return int(10_000 * self.sf)

    @property
    def customers(self) -> int:
        return int(150_000 * self.sf)

    @property
    def orders(self) -> int:
        return int(1_500_000 * self.sf)

    @property
    def lineitems(self) -> int:
        return LINES_PER_ORDER * self.orders


class DeviceQ5:
    """Q5Config's tables in HBM."""

    def __init__(self, ctx, cfg: Q5Config, device: str = "cuda", stream=None):
        import torch

        self.cfg = cfg
        i32 = torch.int32
        k = cfg.seed

        def col(n, sid, lo, hi):
            t = torch.empty(n, dtype=i32, device=device)
            ctx.gen_uniform(t, k, sid, lo, hi, stream=stream)
            return t

        self.r_code = torch.arange(len(REGIONS), dtype=i32, device=device)  # sorted names: rank = index
        self.n_regionkey = torch.tensor(NATION_REGION, dtype=i32, device=device)
        self.n_nationkey = torch.arange(25, dtype=i32, device=device)
        self.s_nationkey = col(cfg.suppliers, 600, 0, 25)
        self.c_nationkey = col(cfg.customers, 601, 0, 25)
        self.o_custkey = col(cfg.orders, 602, 1, cfg.customers + 1)
        y = col(cfg.orders, 603, 1992, 1999)
        m = col(cfg.orders, 604, 1, 13)
        dd = col(cfg.orders, 605, 1, 29)
        self.o_orderdate = (y * 10000 + m * 100 + dd).to(i32)
        del y, m, dd
        L = cfg.lineitems
        self.l_orderkey_new = (torch.arange(L, dtype=torch.int64, device=device)
                               // LINES_PER_ORDER).to(i32)
        self.l_suppkey = col(L, 606, 1, cfg.suppliers + 1)
        self.l_extendedprice = col(L, 607, 90_000, 10_494_951)
        self.l_discount = col(L, 608, 0, 11)

    def fact_bytes(self) -> int:
        return self.cfg.lineitems * 16

    def vectors(self, ctx, region: str, date: int, code_bytes: int = 1, stream=None):
        """Q5's dimension vectors for `region` and orders in [date, date +
        10000) (one year of yyyymmdd), built with general predicates as the
        reference builds them (code/TPCH_Q5/Q5.cpp:915-944): region by name
        -> nation (its key) -> supplier and customer (their nation's key,
        through one hop) -> orders (date range AND the customer's nation,
        through a second hop). Returns (supplier vector, orders vector)."""
        import torch

        from . import olapcore as oc

        dev = self.n_regionkey.device
        codes = {n: i for i, n in enumerate(sorted(REGIONS))}
        i32, dt = torch.int32, torch.int8 if code_bytes == 1 else torch.int32
        region_v = torch.empty(len(REGIONS), dtype=i32, device=dev)
        ctx.build_dimvec_program(oc.DimProgram.of([("in", self.r_code, [codes[region]])]),
                                 len(REGIONS), region_v, stream=stream)
        nation_v = torch.empty(25, dtype=i32, device=dev)
        ctx.build_dimvec_program(oc.DimProgram.of([("hop", self.n_regionkey, region_v, 0)],
                                                  group=oc.GROUP_COLUMN, group_col=self.n_nationkey),
                                 25, nation_v, stream=stream)
        supp_v = torch.empty(self.cfg.suppliers, dtype=dt, device=dev)
        ctx.build_dimvec_program(oc.DimProgram.of([("hop", self.s_nationkey, nation_v, 0)],
                                                  group=oc.GROUP_HOP, group_term=0),
                                 self.cfg.suppliers, supp_v, stream=stream)
        cust_v = torch.empty(self.cfg.customers, dtype=i32, device=dev)
        ctx.build_dimvec_program(oc.DimProgram.of([("hop", self.c_nationkey, nation_v, 0)],
                                                  group=oc.GROUP_HOP, group_term=0),
                                 self.cfg.customers, cust_v, stream=stream)
        order_v = torch.empty(self.cfg.orders, dtype=dt, device=dev)
        ctx.build_dimvec_program(oc.DimProgram.of([("range", self.o_orderdate, date, date + 10000),
                                                   ("and", "hop", self.o_custkey, cust_v, -1)],
                                                  group=oc.GROUP_HOP, group_term=1),
                                 self.cfg.orders, order_v, stream=stream)
        return supp_v, order_v

    def run(self, ctx, supp_v, order_v, out_sum, out_cnt, stream=None):
        """The fused probe over lineitem (Tri_table_join, Q5.cpp:945-978)."""
        ctx.tri_table_join(self.l_suppkey, supp_v, self.l_orderkey_new, order_v,
                           self.l_extendedprice, self.l_discount, 25, out_sum, out_cnt,
                           supp_bias=-1, order_bias=0, stream=stream)
