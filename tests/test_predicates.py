"""Pinning of the oracle's general dimension predicates (no GPU).

- Random programs (1-8 range / IN / hop terms, AND / OR, negation, three
  group kinds, int8 and int32 vectors) against a set-algebra evaluation
  (tests/predicate_cases.py) that shares no code with the oracle.
- A single range term is the plain dimension build (orc_build_dimvec, which
  is pinned against the reference's data generator).
- De Morgan: NOT a OR NOT b is the complement of a AND b.
- The TPC-H Q5 snowflake vectors as the reference's loops write them
  (code/TPCH_Q5/Q5.cpp:921-944: supplier -> nation -> region, orders ->
  customer -> nation with a date range) on a small synthetic schema.
- Dictionary coding: ranks of the distinct strings in byte order."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.predicate_cases import random_program, set_eval


@pytest.mark.parametrize("seed", range(40))
def test_random_programs_match_set_algebra(seed):
    rng = np.random.default_rng(500 + seed)
    rows = int(rng.integers(1, 400))
    pdt = np.int8 if seed % 2 else np.int32
    prow = int(rng.integers(1, 60))
    parent = rng.integers(0, 40, prow).astype(pdt)
    parent[rng.random(prow) < 0.3] = 127 if pdt == np.int8 else -1
    terms = random_program(rng, rows, parent)
    code_bytes = 1 if seed % 3 else 4
    words = [[w for w in t if isinstance(w, str)] for t in terms]
    hops = [i for i, w in enumerate(words) if "hop" in w and "not" not in w]
    kind = ["none", "column", "hop"][seed % 3]
    if kind == "hop" and not hops:
        kind = "column"
    gcol = rng.integers(-5, 300, rows, dtype=np.int32)
    gmod = int(rng.integers(0, 130))
    gterm = hops[0] if hops else 0
    got, bad = po.build_dimvec_program(terms, rows, code_bytes,
                                       group={"none": 0, "column": 1, "hop": 2}[kind],
                                       group_col=gcol, group_mod=gmod, group_term=gterm)
    want = set_eval(terms, rows, kind, gcol, gmod, gterm, code_bytes)
    np.testing.assert_array_equal(got, want)


def test_single_range_is_the_plain_build():
    rng = np.random.default_rng(7)
    f = rng.integers(0, 100, 5000, dtype=np.int32)
    g = rng.integers(0, 1000, 5000, dtype=np.int32)
    for code_bytes, mod in ((4, 0), (4, 77), (1, 100)):
        # the int8 build takes group codes as they are: give it codes below 100
        want = po.build_dimvec(f, g if code_bytes == 4 else g % mod, 10, 60, 1, code_bytes,
                               group_mod=mod if code_bytes == 4 else 0)
        got, _ = po.build_dimvec_program([("range", f, 10, 60)], f.size, code_bytes, group=1,
                                         group_col=g, group_mod=mod)
        np.testing.assert_array_equal(got, want)


def test_de_morgan():
    rng = np.random.default_rng(8)
    a = rng.integers(0, 100, 3000, dtype=np.int32)
    b = rng.integers(0, 100, 3000, dtype=np.int32)
    both, _ = po.build_dimvec_program([("range", a, 20, 70), ("and", "in", b, [1, 5, 9, 50])], a.size)
    either_not, _ = po.build_dimvec_program([("not", "range", a, 20, 70),
                                             ("or", "not", "in", b, [1, 5, 9, 50])], a.size)
    np.testing.assert_array_equal(both == 127, either_not != 127)


def q5_schema(rng):
    """A small TPC-H-shaped snowflake: 5 regions, 25 nations, suppliers,
    customers and orders with 1-based customer keys."""
    r_code = rng.permutation(5).astype(np.int32)      # the region name's dictionary code
    n_regionkey = rng.integers(0, 5, 25, dtype=np.int32)
    s_nationkey = rng.integers(0, 25, 800, dtype=np.int32)
    c_nationkey = rng.integers(0, 25, 3000, dtype=np.int32)
    o_custkey = rng.integers(1, 3001, 20000, dtype=np.int32)
    o_orderdate = rng.integers(19920101, 19981231, 20000, dtype=np.int32)
    return r_code, n_regionkey, s_nationkey, c_nationkey, o_custkey, o_orderdate


def test_q5_snowflake_vectors_as_the_reference_writes_them():
    rng = np.random.default_rng(5)
    r_code, n_regionkey, s_nationkey, c_nationkey, o_custkey, o_orderdate = q5_schema(rng)
    target = int(r_code[2])
    d1, d2 = 19940101, 19950101
    # the reference's loops (Q5.cpp:915-944), transcribed over numpy arrays
    bitmap = (r_code == target).astype(np.int32)
    bitmap_S = np.full(s_nationkey.size, -1, np.int32)
    for s in range(s_nationkey.size):
        if bitmap[n_regionkey[s_nationkey[s]]] == 1:
            bitmap_S[s] = s_nationkey[s]
    bitmap_o = np.full(o_custkey.size, -1, np.int32)
    for o in range(o_custkey.size):
        nk = c_nationkey[o_custkey[o] - 1]
        if bitmap[n_regionkey[nk]] == 1 and d1 <= o_orderdate[o] < d2:
            bitmap_o[o] = nk
    # the same as programs: region -> nation (group: the nation key) ->
    # supplier / customer (group: the nation code the hop finds) -> orders
    region, _ = po.build_dimvec_program([("in", r_code, [target])], 5, 4)
    nation, _ = po.build_dimvec_program([("hop", n_regionkey, region, 0)], 25, 4, group=1,
                                        group_col=np.arange(25, dtype=np.int32))
    supp, _ = po.build_dimvec_program([("hop", s_nationkey, nation, 0)], s_nationkey.size, 4,
                                      group=2, group_term=0)
    cust, _ = po.build_dimvec_program([("hop", c_nationkey, nation, 0)], c_nationkey.size, 4,
                                      group=2, group_term=0)
    orders, _ = po.build_dimvec_program([("range", o_orderdate, d1, d2),
                                         ("and", "hop", o_custkey, cust, -1)],
                                        o_custkey.size, 4, group=2, group_term=1)
    np.testing.assert_array_equal(supp, bitmap_S)
    np.testing.assert_array_equal(orders, bitmap_o)


def test_dict_encode_ranks_distinct_strings():
    from memory_database_amd import olapcore as oc

    names = ["ASIA", "EUROPE", "AFRICA", "ASIA", "AMERICA", "MIDDLE EAST", "EUROPE"]
    codes, nd = oc.dict_encode(names)
    distinct = sorted(set(names))
    assert nd == len(distinct)
    assert list(codes) == [distinct.index(n) for n in names]
    assert oc.dict_code_of(names, codes, "EUROPE") == distinct.index("EUROPE")
    assert oc.dict_code_of(names, codes, "ANTARCTICA") == -1
