"""The reference-signature layer (include/olapcore_compat.hpp) called from
C++ the way the reference's test_multi_compute_operator.cpp calls its
operators: tests/native/compat_harness runs every driver -- OLAPcore_rowwise,
_vwm_dv, _vwm_sv, join_cwm_dv + agg_cwm_dv, join_cwm_sv + agg_cwm_sv and the
NUMA drivers OLAPcore_vwm_dv_numa / _vwm_sv_numa over node shards -- on the
golden fixtures' inputs (written by the reference's own workers). Checked:
the returned int against the fixture's uint32 group vectors summed into an
int as the reference does; the exact int64 results against the oracle; the
per-thread OID / groupID lists and index[] counts against the reference's
row partition; the int16 static gid vector against the fixture's."""
import glob
import os
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "tests", "native", "compat_harness")
FIXTURES = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "olapcore_seed*.npz")))
GROUP_NULL = 32767
DIM_NULL = 127


def legacy_int(u32_vector):
    """int count = 0; count += group_vector[i] (uint32 into int)."""
    s = int(np.asarray(u32_vector, dtype=np.uint64).sum()) & 0xFFFFFFFF
    return s - (1 << 32) if s >= (1 << 31) else s


def expected_lists(d, nthreads):
    """Passing rows and their gids per thread of the reference's partition
    (thread j starts at j * (N / nthreads), the last takes the remainder)."""
    rows = d["m1"].size
    n = int(d["orders"].size)
    ok = np.ones(rows, bool)
    gid = np.zeros(rows, np.int64)
    for i in range(n):
        t = int(d["orders"][i])
        code = d[f"dimvec{t}"][d[f"fk{t}"]].astype(np.int64)
        ok &= code != DIM_NULL
        gid += np.where(code != DIM_NULL, code, 0) * int(d["factor"][i])
    per = rows // nthreads
    out = []
    for j in range(nthreads):
        lo = per * j
        hi = rows if j == nthreads - 1 else lo + per
        r = np.nonzero(ok[lo:hi])[0] + lo
        out.append((r.astype(np.int64), gid[r].astype(np.int16)))
    return out


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p)[:-4] for p in FIXTURES])
@pytest.mark.parametrize("nthreads,nodes", [(1, 1), (3, 2), (7, 3)])
def test_reference_signatures_from_cpp(path, nthreads, nodes, tmp_path, ncpu):
    assert os.path.exists(HARNESS), "build the library (it builds tests/native/compat_harness)"
    d = dict(np.load(path))
    rows = int(d["m1"].size)
    n = int(d["orders"].size)
    G = int(d["group_nums"])
    dim_rows = [int(d[f"dimvec{t}"].size) for t in range(4)]
    with open(tmp_path / "meta.txt", "w") as f:
        f.write(f"{rows} {n} {G} {nthreads} {nodes}\n")
        f.write(" ".join(str(int(x)) for x in d["orders"]) + "\n")
        f.write(" ".join(str(int(x)) for x in d["factor"]) + "\n")
        f.write(" ".join(str(x) for x in dim_rows) + "\n")
    for t in range(4):
        d[f"dimvec{t}"].astype(np.int8).tofile(tmp_path / f"dimvec{t}.bin")
        d[f"fk{t}"].astype(np.int32).tofile(tmp_path / f"fk{t}.bin")
    d["m1"].astype(np.int32).tofile(tmp_path / "m1.bin")
    d["m2"].astype(np.int32).tofile(tmp_path / "m2.bin")
    print(f"compat harness {os.path.basename(path)} rows={rows} G={G} nthreads={nthreads} "
          f"nodes={nodes}", flush=True)
    r = subprocess.run([HARNESS, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    ret = dict(line.split() for line in open(tmp_path / "out.txt"))

    # the oracle's exact results (int64 sum of m1 + m2 and count per slot)
    plan = po.plan(list(d["orders"]), list(d["factor"]), G)
    star = po.Star([d[f"dimvec{t}"] for t in range(4)], [d[f"fk{t}"] for t in range(4)],
                   d["m1"], d["m2"])
    es, ec, bad = po.run("rowwise", plan, star, nthreads=ncpu)
    assert bad == 0
    fixture_of = {"rowwise": "rowwise", "vwm_dv": "vwm_dv", "vwm_sv": "vwm_sv", "cwm_dv": "cwm_dv",
                  "cwm_sv": "cwm_sv", "vwm_dv_numa": "vwm_dv", "vwm_sv_numa": "vwm_sv"}
    for name, fx in fixture_of.items():
        assert int(ret[name]) == legacy_int(d[f"{fx}_sum_u32"]), name
        sums = np.fromfile(tmp_path / f"sums_{name}.bin", dtype=np.int64)
        cnts = np.fromfile(tmp_path / f"counts_{name}.bin", dtype=np.int64)
        np.testing.assert_array_equal(sums, es, err_msg=f"{name} sums")
        np.testing.assert_array_equal(cnts, ec, err_msg=f"{name} counts")
        # the reference's uint32 vectors are the exact results mod 2^32
        np.testing.assert_array_equal(sums.astype(np.uint32), d[f"{fx}_sum_u32"], err_msg=name)
        np.testing.assert_array_equal(cnts.astype(np.uint32), d[f"{fx}_count_u32"], err_msg=name)

    # join_cwm_dv's per-thread lists, in the reference's partition and order
    index = np.fromfile(tmp_path / "cwm_dv_index.bin", dtype=np.int32)
    for j, (oid, gid) in enumerate(expected_lists(d, nthreads)):
        assert index[j] == oid.size, f"thread {j} index"
        np.testing.assert_array_equal(np.fromfile(tmp_path / f"cwm_dv_oid{j}.bin", dtype=np.int64), oid)
        np.testing.assert_array_equal(np.fromfile(tmp_path / f"cwm_dv_gid{j}.bin", dtype=np.int16), gid)

    # join_cwm_sv's int16 static vector: gid or GROUP_NULL per row
    sv = np.fromfile(tmp_path / "cwm_sv_gid.bin", dtype=np.int16)
    np.testing.assert_array_equal(sv, d["cwm_sv_gid"])
    assert (sv == GROUP_NULL).sum() == rows - int(ec.sum())


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p)[:-4] for p in FIXTURES])
def test_raw_uint32_vectors_equal_the_references(ctx, path):
    """Every strategy through the C ABI on device-resident columns, then
    olap_legacy_u32: the reference workers' uint32 group vectors, slot by
    slot (sums of M1 + M2, and counts)."""
    import torch

    from memory_database_amd import olapcore as oc

    d = dict(np.load(path))
    n = int(d["orders"].size)
    G = int(d["group_nums"])
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    cols = oc.Context.columns([dev(d[f"dimvec{t}"]) for t in range(4)],
                              [dev(d[f"fk{t}"]) for t in range(4)], dev(d["m1"]), dev(d["m2"]),
                              rows=int(d["m1"].size))
    plan = oc.plan_from(list(d["orders"])[:n], list(d["factor"])[:n], G)
    q = ctx.query(plan, cols)
    osum = torch.empty(G, dtype=torch.int64, device="cuda")
    ocnt = torch.empty(G, dtype=torch.int64, device="cuda")
    u32 = torch.empty(G, dtype=torch.int32, device="cuda")
    for s in ("rowwise", "vwm_dv", "vwm_sv", "cwm_dv", "cwm_sv"):
        q.run(s, osum, ocnt)
        ctx.legacy_u32(osum, u32)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(u32.cpu().numpy().view(np.uint32), d[f"{s}_sum_u32"], err_msg=s)
        ctx.legacy_u32(ocnt, u32)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(u32.cpu().numpy().view(np.uint32), d[f"{s}_count_u32"], err_msg=s)
    q.close()
