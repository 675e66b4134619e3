"""Write tests/golden/known_totals_sf100.json: the totals the reference
recorded for its GPU OLAPcore runs at SF=100 (the first "Result count" line
of each "Generated data" block in GPU/plot/old.py) paired, in order, with the
flags of the runs in code/GPU_OLAPcore/benchmark.sh, and the five total times
the same block records (one per strategy, in the order of the block's
"Start test" lines). Run in a container that has /root/reference:

    python tests/golden/make_known_totals.py
"""
from __future__ import annotations

import json
import os
import re

REF = os.environ.get("REF", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    runs = open(os.path.join(REF, "code/GPU_OLAPcore/benchmark.sh")).read().strip().splitlines()
    old = open(os.path.join(REF, "GPU/plot/old.py")).read().splitlines()
    totals = []  # (line number, count) of the first result line after each data block
    times = []   # per data block: [(strategy, ms, line number)] in block order
    models = {"row-wise model": "rowwise",
              "Column-wise model and dynamic vector": "cwm_dv",
              "Column-wise model and static vector": "cwm_sv",
              "Vector-wise model and dynamic vector": "vwm_dv",
              "Vector-wise model and static vector": "vwm_sv"}
    want = False
    model = None
    for i, line in enumerate(old, start=1):
        if line.startswith(">>> Generated data"):
            want = True
            times.append([])
        m = re.search(r"Start test GPU OLAPcore using (.*)$", line)
        if m:
            model = models[m.group(1).strip()]
        m = re.search(r"Result count of total selection rate .* is (\d+)/(\d+)", line)
        if m and want:
            totals.append((i, int(m.group(1)), int(m.group(2))))
            want = False
        m = re.search(r"[Tt]otal time: ([0-9.]+)ms", line)
        if m:
            times[-1].append((model, float(m.group(1)), i))
    assert len(totals) == len(runs), (len(totals), len(runs))
    assert all(len(t) == 5 for t in times), [len(t) for t in times]
    cases = []
    for run, (line, count, rows), tm in zip(runs, totals, times):
        flags = dict(re.findall(r"--([a-zA-Z_-]+)=(\S+)", run))
        sq = int(flags.pop("sqlnum"))
        flags.pop("nthreads", None)
        cases.append({"sqlnum": sq, "flags": flags, "count": count, "rows": rows,
                      "source": f"GPU/plot/old.py:{line} (flags: code/GPU_OLAPcore/benchmark.sh)",
                      "published_ms": {k: v for k, v, _ in tm},
                      "published_ms_source": "GPU/plot/old.py:" + ",".join(str(n) for _, _, n in tm)})
    doc = {"about": "Totals the reference recorded for its GPU OLAPcore at SF=100 (600M fact rows, "
                    "M1 = M2 = 5, count = sum of the group vector = 10 x passing rows), with the "
                    "flags of each run. Data: the reference's gen_data with glibc rand() seed 1; "
                    "the byte past each dimension vector (read for FK == D) is 0. published_ms: the "
                    "total time the same run printed per strategy (a single cold run on an unnamed "
                    "NVIDIA GPU).",
           "cases": cases}
    with open(os.path.join(HERE, "known_totals_sf100.json"), "w") as f:
        json.dump(doc, f, indent=1)
    print(f"{len(cases)} cases")


if __name__ == "__main__":
    main()
