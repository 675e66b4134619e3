"""The totals the reference recorded for its ten GPU OLAPcore runs at SF=100
(tests/golden/known_totals_sf100.json), reproduced by the drop-in driver on
the reference's own data generator: glibc rand() seed 1 in gen_data's draw
order, with the byte past each dimension vector set to 0. Every strategy must
give the recorded count.

Generating each SF=100 dataset draws 2.4 G rand() values on one host core
(tens of seconds per query), so two of the ten cases run by default and the
other eight only when OLAP_KNOWN_ANSWERS=1 is set."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURE = os.path.join(ROOT, "tests", "golden", "known_totals_sf100.json")
DRIVER = os.path.join(ROOT, "memory_database_amd", "olapcore_test")
CASES = json.load(open(FIXTURE))["cases"]


def test_fixture_holds_the_ten_recorded_runs():
    assert [c["sqlnum"] for c in CASES] == [21, 22, 23, 31, 32, 33, 34, 41, 42, 43]
    assert all(c["rows"] == 600_000_000 and c["flags"]["SF"] == "100" for c in CASES)
    # count = 10 x passing rows (M1 = M2 = 5)
    assert all(c["count"] % 10 == 0 for c in CASES)


# Q2.1 (the largest group table of the ten, 47,780,820) and Q3.4 (the
# smallest total, 4,050) always run; the other eight with OLAP_KNOWN_ANSWERS=1.
ALWAYS = (21, 34)


def _gate(case):
    if case["sqlnum"] in ALWAYS:
        return pytest.param(case, id=str(case["sqlnum"]))
    return pytest.param(case, id=str(case["sqlnum"]), marks=pytest.mark.skipif(
        os.environ.get("OLAP_KNOWN_ANSWERS") != "1",
        reason="SF=100 host generation takes tens of seconds per query; set OLAP_KNOWN_ANSWERS=1"))


@pytest.mark.gpu
@pytest.mark.parametrize("case", [_gate(c) for c in CASES])
def test_driver_reproduces_recorded_total(case, tmp_path):
    assert os.path.exists(DRIVER), "build the driver first"
    js = tmp_path / "out.jsonl"
    cmd = [DRIVER] + [f"--{k}={v}" for k, v in case["flags"].items()] + [
        f"--sqlnum={case['sqlnum']}", "--gen=reference", "--pad=0", "--strategy=all",
        f"--json={js}", f"--log={tmp_path / 'times.tsv'}"]
    print(" ".join(cmd), flush=True)
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=400)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0
    recs = [json.loads(line) for line in open(js)]
    assert len(recs) == 5
    for r in recs:
        assert r["count"] == case["count"], (r["strategy"], r["count"], case["count"])
        assert r["passing_rows"] * 10 == case["count"]
