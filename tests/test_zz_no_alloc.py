"""The allocation-free row-wise entry point (olap_rowwise_noalloc) against
olap_rowwise, through the C ABI: two calls in a row (the second on a smaller
star, so stale scratch from the first would show), then a star whose packed
dimension vectors exceed the fixed scratch, which must be refused."""
import numpy as np
import pytest

from memory_database_amd import olapcore as oc
from oracle import pyoracle as po
from tests.test_gpu_parity import random_star

pytestmark = pytest.mark.gpu


def _run(ctx, entry, plan_np, dimvecs, fks, m1, m2):
    import torch

    def dev(a):
        return None if a is None else torch.from_numpy(np.ascontiguousarray(a)).cuda()

    n = plan_np.dimvec_nums
    plan = oc.plan_from(list(plan_np.orders)[:n], list(plan_np.factor)[:n], int(plan_np.group_nums))
    cols = oc.Context.columns([dev(d) for d in dimvecs], [dev(f) for f in fks], dev(m1), dev(m2),
                              rows=int(m1.size))
    G = int(plan.group_nums)
    osum = torch.full((G,), -7, dtype=torch.int64, device="cuda")
    ocnt = torch.full((G,), -7, dtype=torch.int64, device="cuda")
    total = ctx.run(entry, plan, cols, osum, ocnt)
    torch.cuda.synchronize()
    return osum.cpu().numpy(), ocnt.cpu().numpy(), total


def test_noalloc_equals_rowwise_twice_then_smaller(ctx):
    rng = np.random.default_rng(2024)
    # a large star (two vectors beyond LDS, G beyond LDS), then a small one
    big = random_star(rng, 4, 3_000_017, code_bytes=4, max_groups=40, big_dim=True,
                      two_measures=True)
    small = random_star(rng, 2, 50_021, code_bytes=1, max_groups=9)
    for name, star in (("big", big), ("small", small)):
        plan_np = star[0]
        print(f"noalloc case {name}: rows={star[3].size} G={int(plan_np.group_nums)}", flush=True)
        ws, wc, wt = _run(ctx, "rowwise", *star)
        gs, gc, gt = _run(ctx, "rowwise_noalloc", *star)
        np.testing.assert_array_equal(gc, wc, err_msg=f"{name}: counts")
        np.testing.assert_array_equal(gs, ws, err_msg=f"{name}: sums")
        assert gt == wt == int(wc.sum())
        # and the oracle, for the small case
        if name == "small":
            es, ec, bad = po.run("rowwise", plan_np, po.Star(*star[1:]), nthreads=4)
            assert bad == 0
            np.testing.assert_array_equal(gc, ec)
            np.testing.assert_array_equal(gs, es)


def test_noalloc_refuses_vectors_beyond_scratch(ctx):
    import torch

    rows = 1000
    D = 3_000_000  # 32-bit codes: 12 MB packed, beyond the 8 MiB scratch
    dv = torch.arange(D, dtype=torch.int32, device="cuda")
    fk = torch.zeros(rows, dtype=torch.int32, device="cuda")
    m1 = torch.ones(rows, dtype=torch.int32, device="cuda")
    plan = oc.plan_from([0], [1], D)
    cols = oc.Context.columns([dv, None, None, None], [fk, None, None, None], m1, rows=rows)
    osum = torch.empty(D, dtype=torch.int64, device="cuda")
    ocnt = torch.empty(D, dtype=torch.int64, device="cuda")
    with pytest.raises(oc.OlapError) as ei:
        ctx.run("rowwise_noalloc", plan, cols, osum, ocnt)
    assert ei.value.status == oc.E_CAPACITY
    assert "scratch" in str(ei.value)
    # the allocating form answers the same query
    assert ctx.run("rowwise", plan, cols, osum, ocnt) == rows
