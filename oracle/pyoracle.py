"""numpy bindings of the CPU oracle (oracle/olap_oracle.c) and of the compiled
reference (oracle/_ref, when it was built).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "build")
REF_DIR = os.path.join(HERE, "_ref")
REF_SRC = "/root/reference"

MAX_DIMS = 4
DIM_NULL_I8 = 127
DIM_NULL_I32 = -1


def build(ref: bool = True, verbose: bool = False) -> None:
    """make the oracle (and oracle/_ref when the reference sources exist)."""
    targets = ["all"]
    if ref and os.path.isdir(REF_SRC):
        targets.append("ref")
    out = None if verbose else subprocess.DEVNULL
    subprocess.run(["make", "-C", HERE, *targets], check=True, stdout=out)


def _has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            return " avx512f" in f.read()
    except OSError:
        return False


class OrcPlan(C.Structure):
    _fields_ = [("dimvec_nums", C.c_int32), ("orders", C.c_int32 * 4),
                ("factor", C.c_int32 * 4), ("group_nums", C.c_int64)]


class OrcColumns(C.Structure):
    _fields_ = [("dimvec", C.c_void_p * 4), ("dim_rows", C.c_int64 * 4),
                ("code_bytes", C.c_int32 * 4), ("fk", C.c_void_p * 4),
                ("m1", C.c_void_p), ("m2", C.c_void_p), ("rows", C.c_int64)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        name = "liboracle_v4.so" if _has_avx512() else "liboracle_v3.so"
        path = os.path.join(BUILD, name)
        if not os.path.exists(path):
            build(ref=False)
        L = C.CDLL(path)
        P = C.c_void_p
        L.orc_mix64.restype = C.c_uint64
        L.orc_mix64.argtypes = [C.c_uint64]
        L.orc_stream_key.restype = C.c_uint64
        L.orc_stream_key.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_gen_uniform_i32.argtypes = [P, C.c_int64, C.c_int64, C.c_uint64, C.c_uint64,
                                          C.c_int32, C.c_int32]
        L.orc_zipf_table.argtypes = [P, C.c_int64, C.c_double]
        L.orc_gen_zipf_i32.argtypes = [P, C.c_int64, C.c_int64, C.c_uint64, C.c_uint64, P,
                                       C.c_int64]
        L.orc_size_of_table.restype = C.c_int64
        L.orc_size_of_table.argtypes = [C.c_int, C.c_double]
        L.orc_srand.argtypes = [C.c_uint]
        L.orc_compat_gen_data.argtypes = [C.c_double] * 5 + [P, P, P, P, P, P]
        L.orc_sort_by_rate.argtypes = [P, P, C.c_int32]
        L.orc_make_plan.restype = C.c_int
        L.orc_make_plan.argtypes = [P, P, P, C.POINTER(OrcPlan)]
        L.orc_build_dimvec_i8.argtypes = [P, P, C.c_int64, C.c_int32, C.c_int32, C.c_int32, P]
        L.orc_build_dimvec_i32.argtypes = [P, P, C.c_int64, C.c_int32, C.c_int32, C.c_int32,
                                           C.c_int32, P]
        for n in ("orc_rowwise", "orc_vwm_dv", "orc_vwm_sv", "orc_cwm_dv", "orc_cwm_sv"):
            f = getattr(L, n)
            f.restype = C.c_int64
            f.argtypes = [C.POINTER(OrcPlan), C.POINTER(OrcColumns), P, P, C.c_int]
        L.orc_join_cwm_sv.argtypes = [C.POINTER(OrcPlan), C.POINTER(OrcColumns), P]
        L.orc_legacy_count.restype = C.c_int32
        L.orc_legacy_count.argtypes = [P, C.c_int64]
        L.orc_starjoin_count.restype = C.c_int64
        L.orc_starjoin_count.argtypes = [C.c_int64, P, P, P, P]
        L.orc_vector_group.restype = C.c_int64
        L.orc_vector_group.argtypes = [C.c_int64, P, P, P, P, C.c_int64, P]
        L.orc_agg_expr.restype = C.c_int64
        L.orc_agg_expr.argtypes = [C.c_int64, P, P, P]
        L.orc_join_build.argtypes = [C.c_int64, P, P, P, C.c_int64, P]
        L.orc_tri_table_join.restype = C.c_int64
        L.orc_tri_table_join.argtypes = [C.c_int64, P, C.c_int32, P, C.c_int64, P, C.c_int32, P,
                                         C.c_int64, P, P, C.c_int32, P, P]
        L.orc_build_dimvec_program.restype = C.c_int64
        L.orc_build_dimvec_program.argtypes = [P, C.c_int64, C.c_int32, P]
        L.orc_join_probe.restype = C.c_int64
        L.orc_join_probe.argtypes = [C.c_int64, P, P, C.c_int64, P]
        L.orc_available_cpus.restype = C.c_int
        _lib = L
    return _lib


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data


# ---------------------------------------------------------------- generators

def stream_key(seed: int, stream: int) -> int:
    return int(lib().orc_stream_key(seed, stream))


def gen_uniform(n: int, seed: int, stream: int, lo: int, hi: int, first_row: int = 0):
    out = np.empty(n, dtype=np.int32)
    lib().orc_gen_uniform_i32(_p(out), n, first_row, seed, stream, lo, hi)
    return out


def zipf_table(d: int, theta: float):
    thr = np.zeros(d, dtype=np.uint64)
    lib().orc_zipf_table(_p(thr), d, theta)
    return thr


def gen_zipf(n: int, seed: int, stream: int, thr: np.ndarray, first_row: int = 0):
    out = np.empty(n, dtype=np.int32)
    thr = np.ascontiguousarray(thr, dtype=np.uint64)
    lib().orc_gen_zipf_i32(_p(out), n, first_row, seed, stream, _p(thr), thr.size)
    return out


def size_of_table(table: int, sf: float) -> int:
    return int(lib().orc_size_of_table(table, sf))


def compat_gen_data(c_sele, s_sele, p_sele, d_sele, sf, bitmap, groups, seed: int = 1):
    """gen_data of the reference (glibc rand()), returns (dimvecs, fks, m1, m2)."""
    L = lib()
    L.orc_srand(seed)
    dv = [np.empty(size_of_table(t, sf), dtype=np.int8) for t in range(4)]
    n = size_of_table(4, sf)
    fk = [np.empty(n, dtype=np.int32) for _ in range(4)]
    m1 = np.empty(n, dtype=np.int32)
    m2 = np.empty(n, dtype=np.int32)
    dvp = (C.c_void_p * 4)(*[_p(a) for a in dv])
    fkp = (C.c_void_p * 4)(*[_p(a) for a in fk])
    bm = (C.c_int * 4)(*bitmap)
    gr = (C.c_int * 4)(*groups)
    L.orc_compat_gen_data(c_sele, s_sele, p_sele, d_sele, sf, bm, gr, dvp, fkp, _p(m1), _p(m2))
    return dv, fk, m1, m2


# ---------------------------------------------------------------- plan

def make_plan(groups, bitmap, sele) -> OrcPlan:
    p = OrcPlan()
    g = (C.c_int32 * 4)(*groups)
    b = (C.c_int32 * 4)(*bitmap)
    s = (C.c_double * 4)(*sele)
    lib().orc_make_plan(g, b, s, C.byref(p))
    return p


def plan(orders, factor, group_nums) -> OrcPlan:
    p = OrcPlan()
    p.dimvec_nums = len(orders)
    for i, o in enumerate(orders):
        p.orders[i] = o
        p.factor[i] = factor[i]
    p.group_nums = group_nums
    return p


def sort_by_rate(sele: Sequence[float], orders: Sequence[int]):
    o = (C.c_int32 * len(orders))(*orders)
    s = (C.c_double * len(sele))(*sele)
    lib().orc_sort_by_rate(s, o, len(orders))
    return list(o)


# ---------------------------------------------------------------- dimension vectors

def build_dimvec(filter_col, group_col, lo, hi, bitmap, code_bytes=1, group_mod=0):
    f = np.ascontiguousarray(filter_col, dtype=np.int32)
    g = np.ascontiguousarray(group_col, dtype=np.int32)
    if code_bytes == 1:
        out = np.empty(f.size, dtype=np.int8)
        lib().orc_build_dimvec_i8(_p(f), _p(g), f.size, lo, hi, bitmap, _p(out))
    else:
        out = np.empty(f.size, dtype=np.int32)
        lib().orc_build_dimvec_i32(_p(f), _p(g), f.size, lo, hi, bitmap, group_mod, _p(out))
    return out


# ---------------------------------------------------------------- strategies

class Star:
    """Holds numpy columns so the pointers in OrcColumns stay valid."""

    def __init__(self, dimvecs, fks, m1, m2=None, rows=None):
        self.dimvecs = [None if d is None else np.ascontiguousarray(d) for d in dimvecs]
        self.fks = [None if f is None else np.ascontiguousarray(f, dtype=np.int32) for f in fks]
        self.m1 = None if m1 is None else np.ascontiguousarray(m1, dtype=np.int32)
        self.m2 = None if m2 is None else np.ascontiguousarray(m2, dtype=np.int32)
        c = OrcColumns()
        for t in range(4):
            d = self.dimvecs[t]
            if d is not None:
                c.dimvec[t] = _p(d)
                c.dim_rows[t] = d.size
                c.code_bytes[t] = d.dtype.itemsize
            else:
                c.code_bytes[t] = 1
            if self.fks[t] is not None:
                c.fk[t] = _p(self.fks[t])
                if rows is None:
                    rows = self.fks[t].size
        c.m1 = _p(self.m1)
        c.m2 = _p(self.m2)
        c.rows = rows if rows is not None else 0
        self.cols = c


def run(strategy: str, p: OrcPlan, star: Star, nthreads: int = 1):
    """Returns (sum int64[G], count int64[G], bad_rows). The strategies read
    M1 for every matching row, as the reference's workers do: a star without
    a measure column is refused (count-only results are checked against a
    run over a zero measure)."""
    if star.m1 is None and star.cols.rows > 0:
        raise ValueError("the oracle's strategies need a measure column")
    G = int(p.group_nums)
    s = np.zeros(G, dtype=np.int64)
    c = np.zeros(G, dtype=np.int64)
    bad = getattr(lib(), "orc_" + strategy)(C.byref(p), C.byref(star.cols), _p(s), _p(c),
                                            nthreads)
    return s, c, int(bad)


def join_cwm_sv(p: OrcPlan, star: Star):
    out = np.empty(star.cols.rows, dtype=np.int32)
    lib().orc_join_cwm_sv(C.byref(p), C.byref(star.cols), _p(out))
    return out


def legacy_count(sums: np.ndarray) -> int:
    a = np.ascontiguousarray(sums, dtype=np.int64)
    return int(lib().orc_legacy_count(_p(a), a.size))


def starjoin_count(dimvecs, fks, want_result=False):
    dv = [np.ascontiguousarray(d, dtype=np.int8) for d in dimvecs]
    fk = [np.ascontiguousarray(f, dtype=np.int32) for f in fks]
    n = fk[0].size
    res = np.empty(n, dtype=np.int32) if want_result else None
    dvp = (C.c_void_p * 4)(*[_p(a) for a in dv])
    dr = (C.c_int64 * 4)(*[a.size for a in dv])
    fkp = (C.c_void_p * 4)(*[_p(a) for a in fk])
    cnt = lib().orc_starjoin_count(n, dvp, dr, fkp, _p(res))
    return int(cnt), res


def vector_group(key1, m1, m2, group_num):
    k = np.ascontiguousarray(key1, dtype=np.int32)
    a = np.ascontiguousarray(m1, dtype=np.int32)
    b = None if m2 is None else np.ascontiguousarray(m2, dtype=np.int32)
    gv = np.zeros(group_num, dtype=np.int64)
    bad = C.c_int64()
    tot = lib().orc_vector_group(k.size, _p(k), _p(a), _p(b), _p(gv), group_num, C.byref(bad))
    return int(tot), gv, bad.value


class _OrcTerm(C.Structure):
    _fields_ = [("kind", C.c_int32), ("combine", C.c_int32), ("negate", C.c_int32),
                ("lo", C.c_int32), ("hi", C.c_int32), ("nvalues", C.c_int32),
                ("parent_code_bytes", C.c_int32), ("key_bias", C.c_int32),
                ("col", C.c_void_p), ("values", C.c_void_p), ("parent", C.c_void_p),
                ("parent_rows", C.c_int64)]


class _OrcProgram(C.Structure):
    _fields_ = [("nterms", C.c_int32), ("group", C.c_int32), ("group_mod", C.c_int32),
                ("group_term", C.c_int32), ("group_col", C.c_void_p), ("term", _OrcTerm * 8)]


def build_dimvec_program(terms, rows, code_bytes=1, group=0, group_col=None, group_mod=0,
                         group_term=0):
    """The oracle's general-predicate dimension vector. terms as
    DimProgram.of takes them, over numpy columns: ("range", col, lo, hi),
    ("in", col, values), ("hop", col, parent, key_bias), optionally preceded
    by "and" / "or" / "not". Returns (vector, rows refused for int8)."""
    p = _OrcProgram()
    p.nterms, p.group, p.group_mod, p.group_term = len(terms), group, group_mod, group_term
    keep = []
    if group_col is not None:
        gc = np.ascontiguousarray(group_col, dtype=np.int32)
        keep.append(gc)
        p.group_col = _p(gc)
    for i, t in enumerate(terms):
        t = list(t)
        comb, neg = 0, 0
        while t[0] in ("and", "or", "not"):
            w = t.pop(0)
            if w == "not":
                neg = 1
            else:
                comb = 0 if w == "and" else 1
        x = p.term[i]
        x.combine, x.negate = comb, neg
        col = np.ascontiguousarray(t[1], dtype=np.int32)
        keep.append(col)
        x.col = _p(col)
        if t[0] == "range":
            x.kind, x.lo, x.hi = 0, int(t[2]), int(t[3])
        elif t[0] == "in":
            v = np.ascontiguousarray(t[2], dtype=np.int32)
            keep.append(v)
            x.kind, x.values, x.nvalues = 1, _p(v), v.size
        else:
            par = np.ascontiguousarray(t[2])
            keep.append(par)
            x.kind, x.parent, x.parent_rows = 2, _p(par), par.size
            x.parent_code_bytes = par.dtype.itemsize
            x.key_bias = int(t[3]) if len(t) > 3 else 0
    out = np.empty(rows, dtype=np.int8 if code_bytes == 1 else np.int32)
    bad = lib().orc_build_dimvec_program(C.byref(p), rows, code_bytes, _p(out))
    return out, int(bad)


def tri_table_join(skey, svec, okey, ovec, price, disc, G, sbias=-1, obias=0):
    """TPC-H Q5's fused probe in fixed point: (sum, count, bad) over int32
    vectors (NULL -1)."""
    a = [np.ascontiguousarray(x, dtype=np.int32) for x in (skey, svec, okey, ovec, price, disc)]
    s = np.zeros(G, np.int64)
    c = np.zeros(G, np.int64)
    bad = lib().orc_tri_table_join(a[0].size, _p(a[0]), sbias, _p(a[1]), a[1].size, _p(a[2]), obias,
                                   _p(a[3]), a[3].size, _p(a[4]), _p(a[5]), G, _p(s), _p(c))
    return s, c, int(bad)


def join_build(key, payload, vector):
    """vecjoin build into `vector` (int32 numpy array, updated in place);
    returns the keys outside the vector."""
    k = np.ascontiguousarray(key, dtype=np.int32)
    v = np.ascontiguousarray(payload, dtype=np.int32)
    assert vector.dtype == np.int32 and vector.flags.c_contiguous
    bad = C.c_int64()
    lib().orc_join_build(k.size, _p(k), _p(v), _p(vector), vector.size, C.byref(bad))
    return bad.value


def join_probe(key, vector):
    """(int64 sum of vector[key], keys outside the vector)."""
    k = np.ascontiguousarray(key, dtype=np.int32)
    v = np.ascontiguousarray(vector, dtype=np.int32)
    bad = C.c_int64()
    s = lib().orc_join_probe(k.size, _p(k), _p(v), v.size, C.byref(bad))
    return int(s), bad.value


def agg_expr(p, q, t):
    p = np.ascontiguousarray(p, dtype=np.int32)
    q = np.ascontiguousarray(q, dtype=np.int32)
    t = np.ascontiguousarray(t, dtype=np.int32)
    return int(lib().orc_agg_expr(p.size, _p(p), _p(q), _p(t)))


def available_cpus() -> int:
    return int(lib().orc_available_cpus())


# ---------------------------------------------------------------- compiled reference

def host_has_avx512() -> bool:
    """The AVX-512 subsets the reference's makefile asks for (F, CD, BW, DQ, VL)."""
    try:
        flags = set(open("/proc/cpuinfo").read().split())
    except OSError:
        return False
    return {"avx512f", "avx512cd", "avx512bw", "avx512dq", "avx512vl"} <= flags


class Reference:
    """The reference's own operator workers, compiled from its sources into
    oracle/_ref. Single-threaded workers over [0, n); uint32 group vectors."""

    def __init__(self, avx512: bool = False):
        """avx512: the build with the reference's AVX-512 flags at -O3
        (used only where this host's CPU has those instructions)."""
        self.ok = False
        self.avx512 = avx512 and host_has_avx512() and os.path.exists(
            os.path.join(REF_DIR, "libref_olapcore_avx512.so"))
        p = os.path.join(REF_DIR, "libref_olapcore_avx512.so" if self.avx512 else "libref_olapcore.so")
        if not os.path.exists(p):
            return
        P = C.c_void_p
        L = C.CDLL(p)
        L.ref_gen_data.restype = C.c_int
        L.ref_gen_data.argtypes = [C.c_double] * 5 + [P, P, P, P, P, P]
        L.ref_size_of_table.restype = C.c_int
        L.ref_size_of_table.argtypes = [C.c_int, C.c_double]
        L.ref_srand.argtypes = [C.c_uint]
        L.ref_sort_by_rate.argtypes = [P, P, C.c_int]
        for n in ("ref_row", "ref_vwm_dv", "ref_vwm_sv", "ref_cwm_dv"):
            getattr(L, n).argtypes = [P, P, P, C.c_int, P, C.c_int, P, P, P]
        L.ref_cwm_sv.argtypes = [P, P, P, C.c_int, P, C.c_int, P, P, P, P]
        for n in ("ref_OLAPcore_vwm_dv", "ref_OLAPcore_rowwise"):
            getattr(L, n).restype = C.c_int
            getattr(L, n).argtypes = [P, P, P, C.c_int, P, C.c_int, C.c_int, P, P, C.c_int]
        self.L = L
        sj = os.path.join(REF_DIR, "libref_starjoin.so")
        gp = os.path.join(REF_DIR, "libref_group.so")
        self.SJ = C.CDLL(sj) if os.path.exists(sj) else None
        if self.SJ is not None:
            self.SJ.ref_starjoin_rowwise.restype = C.c_int
            self.SJ.ref_starjoin_rowwise.argtypes = [C.c_int] + [P] * 8
            self.SJ.ref_starjoin_cwm_sv.restype = C.c_int
            self.SJ.ref_starjoin_cwm_sv.argtypes = [C.c_int] + [P] * 9
        jn = os.path.join(REF_DIR, "libref_join.so")
        self.JN = C.CDLL(jn) if os.path.exists(jn) else None
        if self.JN is not None:
            self.JN.ref_join_build.argtypes = [C.c_longlong, P, P, P]
            self.JN.ref_join_probe.restype = C.c_int
            self.JN.ref_join_probe.argtypes = [C.c_longlong, P, P, C.c_int]
        self.GP = C.CDLL(gp) if os.path.exists(gp) else None
        if self.GP is not None:
            self.GP.ref_vector_group.restype = C.c_longlong
            self.GP.ref_vector_group.argtypes = [C.c_longlong, P, P, P, P, C.c_int]
        self.ok = True

    def gen_data(self, c_sele, s_sele, p_sele, d_sele, sf, bitmap, groups, seed=1):
        L = self.L
        L.ref_srand(seed)
        dv = [np.empty(L.ref_size_of_table(t, sf), dtype=np.int8) for t in range(4)]
        n = L.ref_size_of_table(4, sf)
        fk = [np.empty(n, dtype=np.int32) for _ in range(4)]
        m1 = np.empty(n, dtype=np.int32)
        m2 = np.empty(n, dtype=np.int32)
        dvp = (C.c_void_p * 4)(*[_p(a) for a in dv])
        fkp = (C.c_void_p * 4)(*[_p(a) for a in fk])
        L.ref_gen_data(c_sele, s_sele, p_sele, d_sele, sf, (C.c_int * 4)(*bitmap),
                       (C.c_int * 4)(*groups), dvp, fkp, _p(m1), _p(m2))
        return dv, fk, m1, m2

    def sort_by_rate(self, sele, orders):
        o = (C.c_int * len(orders))(*orders)
        s = (C.c_double * len(sele))(*sele)
        self.L.ref_sort_by_rate(s, o, len(orders))
        return list(o)

    def run(self, strategy: str, dimvecs, fks, orders, factor, group_nums, m1, m2,
            want_gid=False):
        """dimvecs must be int8 and hold a readable byte at every fk value."""
        dv = [np.ascontiguousarray(d, dtype=np.int8) for d in dimvecs]
        fk = [np.ascontiguousarray(f, dtype=np.int32) for f in fks]
        n = int(m1.size)
        dvp = (C.c_void_p * 4)(*[_p(a) for a in dv])
        fkp = (C.c_void_p * 4)(*[_p(a) for a in fk])
        o = (C.c_int * 4)(*list(orders) + [0] * (4 - len(orders)))
        f = (C.c_int * 4)(*list(factor) + [1] * (4 - len(factor)))
        gv = np.zeros(group_nums, dtype=np.uint32)
        a = np.ascontiguousarray(m1, dtype=np.int32)
        b = np.ascontiguousarray(m2, dtype=np.int32)
        if strategy == "cwm_sv":
            gid = np.zeros(n, dtype=np.int16)
            self.L.ref_cwm_sv(dvp, fkp, o, len(orders), f, n, _p(a), _p(b), _p(gv), _p(gid))
            return (gv, gid) if want_gid else gv
        fn = {"rowwise": self.L.ref_row, "vwm_dv": self.L.ref_vwm_dv,
              "vwm_sv": self.L.ref_vwm_sv, "cwm_dv": self.L.ref_cwm_dv}[strategy]
        fn(dvp, fkp, o, len(orders), f, n, _p(a), _p(b), _p(gv))
        return gv

    def olapcore_rowwise(self, dimvecs, fks, orders, factor, group_nums, m1, m2, nthreads):
        return self._driver("ref_OLAPcore_rowwise", dimvecs, fks, orders, factor, group_nums, m1,
                            m2, nthreads)

    def olapcore_vwm_dv(self, dimvecs, fks, orders, factor, group_nums, m1, m2, nthreads):
        return self._driver("ref_OLAPcore_vwm_dv", dimvecs, fks, orders, factor, group_nums, m1,
                            m2, nthreads)

    def _driver(self, name, dimvecs, fks, orders, factor, group_nums, m1, m2, nthreads):
        dv = [np.ascontiguousarray(d, dtype=np.int8) for d in dimvecs]
        dvp = (C.c_void_p * 4)(*[_p(a) for a in dv])
        fkp = (C.c_void_p * 4)(*[_p(a) for a in fks])
        o = (C.c_int * 4)(*list(orders) + [0] * (4 - len(orders)))
        f = (C.c_int * 4)(*list(factor) + [1] * (4 - len(factor)))
        return int(getattr(self.L, name)(dvp, fkp, o, len(orders), f, int(m1.size),
                                         int(group_nums), _p(m1), _p(m2), nthreads))
