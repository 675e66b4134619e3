"""The sub-step operators at the reference's own sizes.

- Dense-vector join (code/join/test_join_option.cpp:30-82): build + probe
  against the golden fixtures written by the reference's functions, against
  the oracle over LDS-resident and HBM-resident vectors, repeated and
  out-of-range keys and unaligned key columns; and at the reference's probe
  size (2^30 probe rows) through closed forms: payload 1 sums to the probe
  row count, payload = key sums to the probe keys' sum.
- vector_group (code/group/test_group_option.cpp:44-60) at 2^20, 2^24 and
  2^26 groups, where the group table leaves LDS: against the oracle on 40M
  rows, and on the reference's 600M rows against torch's own index_add_
  of the same int64 values."""
import glob
import os

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _say(*a):
    print(*a, flush=True)


def _dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "join_*.npz"))),
                         ids=lambda p: os.path.basename(p)[:-4])
def test_join_matches_golden(ctx, path):
    import torch

    z = np.load(path)
    vec = torch.zeros(z["vector"].size, dtype=torch.int32, device="cuda")
    assert ctx.join_build(_dev(z["r_key"]), _dev(z["r_payload"]), vec) == 0
    np.testing.assert_array_equal(vec.cpu().numpy(), z["vector"])
    total, bad = ctx.join_probe(_dev(z["s_key"]), vec)
    assert bad == 0
    assert np.int64(total).astype(np.int32) == np.int32(z["count"])


@pytest.mark.parametrize("vlen,rrows,srows,dup,offset", [
    (1 << 10, 1 << 10, 1_000_003, False, 0),      # LDS-resident vector
    (32768, 50_000, 2_000_001, True, 1),          # 128 KiB in LDS, repeated keys, unaligned keys
    (40960, 40960, 3_000_001, False, 0),          # the whole LDS (160 KiB), 1024 threads
    (40961, 40961, 3_000_001, False, 0),          # one entry past it: gathered through L2
    (1 << 22, 1 << 22, 8_000_000, False, 0),      # HBM-resident
    (3_000_017, 2_000_000, 5_000_005, True, 3),   # HBM, repeated, unaligned
])
def test_join_matches_oracle(ctx, vlen, rrows, srows, dup, offset):
    import torch

    rng = np.random.default_rng(vlen + rrows)
    _say(f"join vlen={vlen} build={rrows} probe={srows} dup={dup} offset={offset}")
    if dup:
        rkey = rng.integers(-5, vlen + 5, rrows, dtype=np.int32)
    else:
        rkey = rng.permutation(np.arange(rrows, dtype=np.int32) % vlen)
    rpay = rng.integers(-(2 ** 31), 2 ** 31 - 1, rrows, dtype=np.int32)
    skey = rng.integers(-3, vlen + 3, srows, dtype=np.int32)
    init = rng.integers(-100, 100, vlen, dtype=np.int32)  # entries no key names keep these
    want = init.copy()
    wbad = po.join_build(rkey, rpay, want)
    wtot, wpbad = po.join_probe(skey, want)
    vec = _dev(init)
    buf = torch.empty(srows + offset, dtype=torch.int32, device="cuda")
    buf[offset:] = _dev(skey)
    assert ctx.join_build(_dev(rkey), _dev(rpay), vec) == wbad
    np.testing.assert_array_equal(vec.cpu().numpy(), want)
    assert ctx.join_probe(buf[offset:], vec) == (wtot, wpbad)


def test_join_at_reference_probe_size(ctx):
    """2^20 build rows with unique keys 1..2^20 (the reference's generator's
    key set), 2^30 probe keys uniform over them."""
    import torch

    maxid, srows = 1 << 20, 1 << 30
    rng = np.random.default_rng(41)
    rkey = _dev(rng.permutation(np.arange(1, maxid + 1, dtype=np.int32)))
    skey = torch.empty(srows, dtype=torch.int32, device="cuda")
    ctx.gen_uniform(skey, 41, 7, 1, maxid + 1)
    vec = torch.zeros(maxid + 1, dtype=torch.int32, device="cuda")
    assert ctx.join_build(rkey, torch.ones(maxid, dtype=torch.int32, device="cuda"), vec) == 0
    assert ctx.join_probe(skey, vec) == (srows, 0)
    assert ctx.join_build(rkey, rkey, vec) == 0
    total, bad = ctx.join_probe(skey, vec)
    assert bad == 0 and total == int(skey.to(torch.int64).sum().item())
    _say(f"join 2^30 probes: sum {total}")


@pytest.mark.parametrize("log2g", [20, 24, 26])
def test_vector_group_beyond_lds_matches_oracle(ctx, log2g):
    import torch

    G = 1 << log2g
    rows = 40_000_000
    rng = np.random.default_rng(log2g)
    _say(f"vector_group G=2^{log2g} rows={rows}")
    key = rng.integers(1, G + 1, rows, dtype=np.int32)
    key[::1000] = 0  # outside [1, G]: skipped and counted
    m1 = rng.integers(-(2 ** 31), 2 ** 31 - 1, rows, dtype=np.int32)
    m2 = rng.integers(0, 1000, rows, dtype=np.int32)
    tot, gv, bad = po.vector_group(key, m1, m2, G)
    out = torch.empty(G, dtype=torch.int64, device="cuda")
    gt, gbad = ctx.vector_group(_dev(key), _dev(m1), _dev(m2), out)
    assert (gt, gbad) == (tot, bad)
    np.testing.assert_array_equal(out.cpu().numpy(), gv)


@pytest.mark.parametrize("log2g,m2,offset", [(17, False, 0), (21, False, 1), (21, True, 3), (27, False, 0)])
def test_vector_group_forms(ctx, log2g, m2, offset):
    """The staged form with one measure (8-byte records) and with two, on
    aligned and unaligned columns (scalar loads), and 2^27 groups, whose
    partition state does not fit LDS: global atomics."""
    import torch

    G = 1 << log2g
    rows = 3_000_007
    rng = np.random.default_rng(log2g + offset)
    key = rng.integers(0, G + 2, rows, dtype=np.int32)   # 0 and G + 1 are outside [1, G]
    a = rng.integers(-(2 ** 31), 2 ** 31 - 1, rows, dtype=np.int32)
    b = rng.integers(0, 1000, rows, dtype=np.int32) if m2 else None
    tot, gv, bad = po.vector_group(key, a, b, G)

    def up(x):
        if x is None:
            return None
        buf = torch.empty(x.size + offset, dtype=torch.int32, device="cuda")
        buf[offset:] = torch.from_numpy(x).cuda()
        return buf[offset:]

    _say(f"vector_group G=2^{log2g} m2={m2} offset={offset}")
    out = torch.empty(G, dtype=torch.int64, device="cuda")
    gt, gbad = ctx.vector_group(up(key), up(a), up(b), out)
    assert (gt, gbad) == (tot, bad)
    np.testing.assert_array_equal(out.cpu().numpy(), gv)


@pytest.mark.parametrize("log2g", [20, 26])
def test_vector_group_at_reference_size(ctx, log2g):
    """600M rows (the reference's DATA_NUM_BASE x SF=100) against torch's
    index_add_ of the same int64 values."""
    import torch

    G = 1 << log2g
    rows = 600_000_000
    key = torch.empty(rows, dtype=torch.int32, device="cuda")
    m1 = torch.empty(rows, dtype=torch.int32, device="cuda")
    m2 = torch.empty(rows, dtype=torch.int32, device="cuda")
    ctx.gen_uniform(key, 5, log2g, 1, G + 1)
    ctx.gen_uniform(m1, 5, 100 + log2g, -(2 ** 30), 2 ** 30)
    ctx.gen_uniform(m2, 5, 200 + log2g, 0, 1000)
    out = torch.empty(G, dtype=torch.int64, device="cuda")
    gt, gbad = ctx.vector_group(key, m1, m2, out)
    want = torch.zeros(G, dtype=torch.int64, device="cuda")
    want.index_add_(0, (key - 1).to(torch.int64), m1.to(torch.int64) + m2.to(torch.int64))
    assert gbad == 0
    assert torch.equal(out, want)
    assert gt == int(want.sum().item())
    _say(f"vector_group 600M rows G=2^{log2g}: total {gt}")
