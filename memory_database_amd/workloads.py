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
