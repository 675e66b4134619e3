"""Host copies of the benchmark stars, generated with the oracle's
counter-based generator (same (seed, stream, row) definition as the device
generator), plus a brute-force nested-loop star join for tiny inputs."""
from __future__ import annotations

import numpy as np

from oracle import pyoracle as po


def cpu_star(cfg, first_row: int = 0, nrows: int | None = None, code_bytes: int = 1):
    """(dimvecs, filters, groups, fks, m1, m2) by table slot for rows
    [first_row, first_row + nrows) of the fact table."""
    n = cfg.rows - first_row if nrows is None else nrows
    dimvecs, filters, groups, fks = [None] * 4, [None] * 4, [None] * 4, [None] * 4
    for d in cfg.dims:
        f = po.gen_uniform(d.rows, cfg.seed, 100 + d.slot, 0, 100)
        g = po.gen_uniform(d.rows, cfg.seed, 200 + d.slot, 0, d.groups)
        dimvecs[d.slot] = po.build_dimvec(f, g, d.pred_lo, d.pred_hi, d.bitmap, code_bytes,
                                          group_mod=0)
        filters[d.slot], groups[d.slot] = f, g
        if cfg.zipf_theta > 0.0:
            thr = po.zipf_table(d.rows, cfg.zipf_theta)
            fks[d.slot] = po.gen_zipf(n, cfg.seed, 300 + d.slot, thr, first_row)
        else:
            fks[d.slot] = po.gen_uniform(n, cfg.seed, 300 + d.slot, 0, d.rows, first_row)
    m1 = po.gen_uniform(n, cfg.seed, 400, cfg.m_lo, cfg.m_hi + 1, first_row)
    m2 = po.gen_uniform(n, cfg.seed, 401, cfg.m_lo, cfg.m_hi + 1, first_row) if cfg.two_measures else None
    return dimvecs, filters, groups, fks, m1, m2


def brute_force(filters, groups, preds, bitmaps, fks, m1, m2, orders, factor, group_nums):
    """Nested-loop star join straight from filter/group columns: for every
    fact row and every probed table, scan the dimension rows for the one whose
    key equals the foreign key. Independent of dimension vectors and of every
    strategy. Returns {gid: (sum, count)} and the number of out-of-range gids."""
    out = {}
    bad = 0
    n = len(m1)
    for i in range(n):
        gid = 0
        ok = True
        for s, t in enumerate(orders):
            key = int(fks[t][i])
            match = None
            for r in range(len(filters[t])):
                if r == key:
                    match = r
                    break
            if match is None:
                ok = False
                break
            lo, hi = preds[t]
            if not (lo <= int(filters[t][match]) < hi):
                ok = False
                break
            code = int(groups[t][match]) if bitmaps[t] else 0
            gid += code * factor[s]
        if not ok:
            continue
        if gid < 0 or gid >= group_nums:
            bad += 1
            continue
        m = int(m1[i]) + (int(m2[i]) if m2 is not None else 0)
        s_, c_ = out.get(gid, (0, 0))
        out[gid] = (s_ + m, c_ + 1)
    return out, bad


def dense(result: dict, group_nums: int):
    s = np.zeros(group_nums, dtype=np.int64)
    c = np.zeros(group_nums, dtype=np.int64)
    for g, (a, b) in result.items():
        s[g] = a
        c[g] = b
    return s, c
