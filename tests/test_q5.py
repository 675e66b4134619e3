"""Pinning of the oracle's TPC-H Q5 fused probe (no GPU): against a numpy
formulation (fancy indexing and np.add.at, exact int64), and against the
reference's own double arithmetic (l_extendedprice * (1 - l_discount),
code/TPCH_Q5/Q5.cpp:945-978) within double rounding."""
import numpy as np
import pytest

from memory_database_amd import workloads as wl
from oracle import pyoracle as po
from tests.q5_host import host_q5, reference_vectors


@pytest.mark.parametrize("region,date", [(2, 19940101), (0, 19960101), (4, 19921231)])
def test_q5_oracle_matches_numpy_and_double(region, date):
    d = host_q5(wl.Q5Config(0.05))
    sv, ov = reference_vectors(d, region, date)
    s, c, bad = po.tri_table_join(d["l_suppkey"], sv, d["l_orderkey_new"], ov,
                                  d["l_extendedprice"], d["l_discount"], 25)
    assert bad == 0
    scode = sv[d["l_suppkey"] - 1]
    ocode = ov[d["l_orderkey_new"]]
    m = (scode != -1) & (scode == ocode)
    ws = np.zeros(25, np.int64)
    wc = np.zeros(25, np.int64)
    np.add.at(ws, scode[m], d["l_extendedprice"][m].astype(np.int64) * (100 - d["l_discount"][m]))
    np.add.at(wc, scode[m], 1)
    np.testing.assert_array_equal(s, ws)
    np.testing.assert_array_equal(c, wc)
    assert c.sum() > 0 and (c > 0).sum() == 5  # the region's five nations
    # the reference's double revenue
    grp = np.zeros(25)
    price = d["l_extendedprice"][m] / 100.0
    disc = d["l_discount"][m] / 100.0
    np.add.at(grp, scode[m], price * (1 - disc))
    np.testing.assert_allclose(s / 1e4, grp, rtol=1e-11)


def test_q5_keys_outside_vectors_find_no_row():
    sv = np.array([3, -1, 7], np.int32)
    ov = np.array([3, 7, 7, -1], np.int32)
    skey = np.array([1, 2, 3, 3, 4, 0, 1], np.int32)   # 1-based; 4 and 0 fall outside
    okey = np.array([0, 0, 1, 2, 0, 0, 9], np.int32)
    price = np.full(7, 100, np.int32)
    disc = np.array([0, 0, 5, 10, 0, 0, 0], np.int32)
    s, c, bad = po.tri_table_join(skey, sv, okey, ov, price, disc, 8)
    assert bad == 0
    assert c[3] == 1 and s[3] == 100 * 100          # row 0
    assert c[7] == 2 and s[7] == 100 * 95 + 100 * 90  # rows 2 and 3
    assert c.sum() == 3
