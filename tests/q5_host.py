"""Host copies of the TPC-H-shaped Q5 tables (memory_database_amd.workloads
.Q5Config) drawn from the oracle's generator, and the reference's Q5 vector
loops (code/TPCH_Q5/Q5.cpp:915-944) transcribed over them."""
import numpy as np

from memory_database_amd import workloads as wl
from oracle import pyoracle as po


def host_q5(cfg):
    k = cfg.seed
    d = {"n_regionkey": np.array(wl.NATION_REGION, np.int32)}
    d["s_nationkey"] = po.gen_uniform(cfg.suppliers, k, 600, 0, 25)
    d["c_nationkey"] = po.gen_uniform(cfg.customers, k, 601, 0, 25)
    d["o_custkey"] = po.gen_uniform(cfg.orders, k, 602, 1, cfg.customers + 1)
    y = po.gen_uniform(cfg.orders, k, 603, 1992, 1999).astype(np.int64)
    m = po.gen_uniform(cfg.orders, k, 604, 1, 13).astype(np.int64)
    dd = po.gen_uniform(cfg.orders, k, 605, 1, 29).astype(np.int64)
    d["o_orderdate"] = (y * 10000 + m * 100 + dd).astype(np.int32)
    L = cfg.lineitems
    d["l_orderkey_new"] = (np.arange(L, dtype=np.int64) // wl.LINES_PER_ORDER).astype(np.int32)
    d["l_suppkey"] = po.gen_uniform(L, k, 606, 1, cfg.suppliers + 1)
    d["l_extendedprice"] = po.gen_uniform(L, k, 607, 90_000, 10_494_951)
    d["l_discount"] = po.gen_uniform(L, k, 608, 0, 11)
    return d


def reference_vectors(d, region: int, date: int):
    """set_bitmap, DimVec_S and DimVec_o as the reference's loops write them
    (Q5.cpp:915-944), vectorised over numpy: -1 where the row fails."""
    bitmap = (np.arange(5) == region).astype(np.int32)
    s_ok = bitmap[d["n_regionkey"][d["s_nationkey"]]] == 1
    bitmap_S = np.where(s_ok, d["s_nationkey"], -1).astype(np.int32)
    nk = d["c_nationkey"][d["o_custkey"] - 1]
    o_ok = (bitmap[d["n_regionkey"][nk]] == 1) & (d["o_orderdate"] >= date) & (d["o_orderdate"] < date + 10000)
    bitmap_o = np.where(o_ok, nk, -1).astype(np.int32)
    return bitmap_S, bitmap_o
