"""Write the golden fixtures of the star-join path from the reference's own
operator code, compiled from its sources into oracle/_ref (oracle/Makefile,
`make ref`). Run in a container that has /root/reference:

    python tests/golden/make_golden.py

Each fixture (tests/golden/olapcore_*.npz) holds the inputs -- int8 dimension
vectors (padded with one NULL byte so the reference's inclusive FK range stays
in bounds), foreign keys, two measure columns, the plan -- and the outputs of
the reference's single-thread workers OLAPcore_row_thread,
OLAPcore_vwm_dv_thread, OLAPcore_vwm_sv_thread, join/agg_cwm_dv_thread and
join/agg_cwm_sv_thread: their uint32 group vectors for the measures and for
unit measures (= counts), and the int16 static gid vector of join_cwm_sv.
starjoin_*.npz, join_*.npz and group_*.npz hold the sub-step operators'
results (`python tests/golden/make_golden.py join` writes only join_*.npz).
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyoracle as po  # noqa: E402

STRATS = ["rowwise", "vwm_dv", "vwm_sv", "cwm_dv", "cwm_sv"]


def star_case(seed: int, ndims: int, rows: int):
    rng = np.random.default_rng(seed)
    slots = list(rng.permutation(4)[:ndims])
    groups = [0] * 4
    bitmap = [0] * 4
    sele = [0.0] * 4
    dimvecs = []
    fks = []
    for t in range(4):
        if t in slots:
            D = int(rng.integers(1, 3000))
            g = int(rng.integers(1, 12))
            bm = int(rng.integers(0, 2)) if g == 1 else 1
            groups[t], bitmap[t] = g, bm
            sele[t] = float(rng.choice([0.05, 0.2, 0.5, 0.9, 1.0]))
            codes = rng.integers(0, g, D).astype(np.int8) if bm else np.zeros(D, np.int8)
            codes[rng.random(D) > sele[t]] = 127
        else:
            D = 1
            codes = np.full(1, 127, np.int8)
        # inclusive FK range [0, D] as gen_data draws it; byte D is a NULL pad
        dimvecs.append(np.concatenate([codes, np.array([127], np.int8)]))
        fks.append(rng.integers(0, D + 1, rows, dtype=np.int32))
    plan = po.make_plan(groups, bitmap, sele)
    m1 = rng.integers(1, 50, rows, dtype=np.int32)
    m2 = rng.integers(0, 50, rows, dtype=np.int32)
    return dimvecs, fks, m1, m2, plan, groups, bitmap, sele


def join_cases(ref):
    """vecjoin build + probe (code/join/test_join_option.cpp:30-36, :58-82)
    through the reference's own functions: unique keys 1..maxid shuffled as
    its generator makes them (gendata_util.hpp:259-300, payload 1), the same
    with random payloads, and a build relation with repeated keys (the later
    row wins)."""
    rng = np.random.default_rng(23)
    for name, maxid, s_rows, payload_kind, dup in (("unique_ones", 1000, 50_000, "ones", False),
                                                  ("unique_rand", 4096, 40_000, "rand", False),
                                                  ("repeated", 777, 30_000, "rand", True)):
        if dup:
            rkey = rng.integers(1, maxid + 1, 3 * maxid, dtype=np.int32)
        else:
            rkey = rng.permutation(np.arange(1, maxid + 1, dtype=np.int32))
        rpay = (np.ones(rkey.size, np.int32) if payload_kind == "ones"
                else rng.integers(-1000, 1000, rkey.size, dtype=np.int32))
        skey = rng.integers(1, maxid + 1, s_rows, dtype=np.int32)
        vec = np.zeros(maxid + 1, np.int32)  # the reference leaves unnamed entries as they are
        ref.JN.ref_join_build(rkey.size, rkey.ctypes.data, rpay.ctypes.data, vec.ctypes.data)
        cnt = ref.JN.ref_join_probe(skey.size, skey.ctypes.data, vec.ctypes.data, 0)
        cnt_pf = ref.JN.ref_join_probe(skey.size, skey.ctypes.data, vec.ctypes.data, 1)
        path = os.path.join(HERE, f"join_{name}.npz")
        np.savez_compressed(path, r_key=rkey, r_payload=rpay, s_key=skey, vector=vec,
                            count=np.int64(cnt), count_prefetch=np.int64(cnt_pf))
        print("wrote", path)


def main():
    ref = po.Reference()
    if not ref.ok:
        raise SystemExit("oracle/_ref is not built (make -C oracle ref)")
    if sys.argv[1:] == ["join"]:
        return join_cases(ref)
    cases = [(11, 1, 5000), (12, 2, 7001), (13, 3, 6000), (14, 4, 9000), (15, 4, 3333),
             (16, 2, 1025)]
    for seed, nd, rows in cases:
        dv, fk, m1, m2, plan, groups, bitmap, sele = star_case(seed, nd, rows)
        orders = list(plan.orders)[: plan.dimvec_nums]
        factor = list(plan.factor)[: plan.dimvec_nums]
        G = int(plan.group_nums)
        out = {}
        ones = np.ones(rows, np.int32)
        zeros = np.zeros(rows, np.int32)
        for s in STRATS:
            r = ref.run(s, dv, fk, orders, factor, G, m1, m2, want_gid=(s == "cwm_sv"))
            if s == "cwm_sv":
                r, gid = r
                out["cwm_sv_gid"] = gid
            out[f"{s}_sum_u32"] = r
            c = ref.run(s, dv, fk, orders, factor, G, ones, zeros)
            out[f"{s}_count_u32"] = c[0] if isinstance(c, tuple) else c
        path = os.path.join(HERE, f"olapcore_seed{seed}_{nd}dims.npz")
        np.savez_compressed(path, dimvec0=dv[0], dimvec1=dv[1], dimvec2=dv[2], dimvec3=dv[3],
                            fk0=fk[0], fk1=fk[1], fk2=fk[2], fk3=fk[3], m1=m1, m2=m2,
                            orders=np.array(orders, np.int32), factor=np.array(factor, np.int32),
                            group_nums=np.int64(G), groups=np.array(groups, np.int32),
                            bitmap=np.array(bitmap, np.int32), sele=np.array(sele), **out)
        print("wrote", path)

    # star join (count only) sub-step, code/starjoin/test_starjoin_option.cpp:37-121
    if ref.SJ is not None:
        rng = np.random.default_rng(21)
        rows = 20000
        dv, fk = [], []
        for D in (300, 20, 2000, 255):
            d = rng.integers(0, 16, D).astype(np.int8)
            d[rng.random(D) < 0.3] = 127
            dv.append(np.concatenate([d, np.array([127], np.int8)]))
            fk.append(rng.integers(0, D + 1, rows, dtype=np.int32))
        P = lambda a: a.ctypes.data  # noqa: E731
        cnt = ref.SJ.ref_starjoin_rowwise(rows, *[P(d) for d in dv], *[P(f) for f in fk])
        res = np.zeros(rows, np.int32)
        cnt2 = ref.SJ.ref_starjoin_cwm_sv(rows, *[P(d) for d in dv], *[P(f) for f in fk], P(res))
        path = os.path.join(HERE, "starjoin_seed21.npz")
        np.savez_compressed(path, dimvec0=dv[0], dimvec1=dv[1], dimvec2=dv[2], dimvec3=dv[3],
                            fk0=fk[0], fk1=fk[1], fk2=fk[2], fk3=fk[3], count=np.int64(cnt),
                            count_cwm_sv=np.int64(cnt2), result=res)
        print("wrote", path)

    if ref.JN is not None:
        join_cases(ref)

    # dense-vector group-by sub-step, code/group/test_group_option.cpp:44-60
    if ref.GP is not None:
        rng = np.random.default_rng(22)
        for G in (32, 1024):
            rows = 16000
            key = rng.integers(1, G + 1, rows, dtype=np.int32)
            m1 = rng.integers(0, 1000, rows, dtype=np.int32)
            m2 = rng.integers(0, 1000, rows, dtype=np.int32)
            gv = np.zeros(G, np.int32)
            tot = ref.GP.ref_vector_group(rows, key.ctypes.data, m1.ctypes.data, m2.ctypes.data,
                                          gv.ctypes.data, G)
            path = os.path.join(HERE, f"group_G{G}.npz")
            np.savez_compressed(path, key=key, m1=m1, m2=m2, group_vector=gv, total=np.int64(tot))
            print("wrote", path)


if __name__ == "__main__":
    main()
