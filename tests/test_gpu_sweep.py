"""The sweep of the 4-dimension star on the GPU against the oracle: Zipf
skew 0 / 0.5 / 1 of every foreign key, 10^2 / 10^4 / 10^6 groups and 1% / 10%
/ 50% selectivity per dimension, at 2M fact rows (bit-exact per slot), and at
the full 600M rows through size-independent properties (strategies agree;
the counts add up to the star join's passing rows)."""
import numpy as np
import pytest

from memory_database_amd import olapcore as oc
from memory_database_amd import workloads as wl
from oracle import pyoracle as po
from tests.cpu_star import cpu_star

pytestmark = pytest.mark.gpu

ROWS = 2_000_003
CASES = wl.sweep_configs(ROWS)


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("cfg", CASES, ids=[c.name.split("M_", 1)[1] for c in CASES])
def test_sweep_matches_oracle(ctx, ncpu, cfg):
    import torch

    _say(f"{cfg.name}")
    ds = wl.DeviceStar(ctx, cfg)
    dimvecs, _, _, fks, m1, m2 = cpu_star(cfg)
    plan = ds.plan
    p = po.plan(list(plan.orders)[: plan.dimvec_nums], list(plan.factor)[: plan.dimvec_nums],
                int(plan.group_nums))
    es, ec, bad = po.run("rowwise", p, po.Star(dimvecs, fks, m1, m2), nthreads=ncpu)
    assert bad == 0
    G = int(plan.group_nums)
    q = ctx.query(plan, ds.cols)
    strategies = ["rowwise", "vwm_sv"] if G > 10_000 else list(oc.STRATEGIES)
    for s in strategies:
        osum = torch.full((G,), -7, dtype=torch.int64, device="cuda")
        ocnt = torch.full((G,), -7, dtype=torch.int64, device="cuda")
        q.run(s, osum, ocnt)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(ocnt.cpu().numpy(), ec, err_msg=f"{s} counts")
        np.testing.assert_array_equal(osum.cpu().numpy(), es, err_msg=f"{s} sums")
    # the row-wise scan with the other aggregation layout gives the same slots
    if G > 10_000:
        for layout in (0, 1):
            q.set_aggregation(layout)
            osum = torch.full((G,), -7, dtype=torch.int64, device="cuda")
            ocnt = torch.full((G,), -7, dtype=torch.int64, device="cuda")
            q.run("rowwise", osum, ocnt)
            torch.cuda.synchronize()
            np.testing.assert_array_equal(ocnt.cpu().numpy(), ec, err_msg=f"layout {layout} counts")
            np.testing.assert_array_equal(osum.cpu().numpy(), es, err_msg=f"layout {layout} sums")
    q.close()


@pytest.mark.parametrize("theta,G,sel", [(1.0, 1_000_000, 50), (0.5, 10_000, 10)])
def test_sweep_full_size_properties(ctx, theta, G, sel):
    """600M rows: every strategy gives the same slots, and the counts add up
    to the rows the star join (olap_starjoin, no grouping) lets through."""
    import torch

    cfg = [c for c in wl.sweep_configs(600_000_000, thetas=(theta,), groups=(G,), seles=(sel,))][0]
    _say(f"{cfg.name}")
    ds = wl.DeviceStar(ctx, cfg)
    q = ctx.query(ds.plan, ds.cols)
    ref = None
    for s in ["rowwise", "vwm_sv", "vwm_dv", "cwm_dv", "cwm_sv"]:
        osum = torch.full((G,), -7, dtype=torch.int64, device="cuda")
        ocnt = torch.full((G,), -7, dtype=torch.int64, device="cuda")
        q.run(s, osum, ocnt)
        torch.cuda.synchronize()
        got = (osum.cpu().numpy(), ocnt.cpu().numpy())
        if ref is None:
            ref = got
        else:
            np.testing.assert_array_equal(got[1], ref[1], err_msg=f"{s} counts")
            np.testing.assert_array_equal(got[0], ref[0], err_msg=f"{s} sums")
    q.close()
    dvs = [ds.dimvecs[t] for t in range(4)]
    passing = ctx.starjoin("rowwise", dvs, ds.fks)
    _say(f"  passing rows {passing}, counted {int(ref[1].sum())}")
    assert int(ref[1].sum()) == passing
    assert passing > 0
    # measures are U[1000, 1000000]: every group's sum lies within its count's bounds
    c, s = ref[1], ref[0]
    assert np.all(s >= 1000 * c) and np.all(s <= 1_000_000 * c)
