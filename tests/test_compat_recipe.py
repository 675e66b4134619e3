"""The drop-in recipe for the reference's operator file links: take
code/OLAPcore/test_multi_compute_operator.cpp as it lies in the reference,
drop its operator definitions (lines 685-1437: agg_cwm_dv through
OLAPcore_vwm_sv_numa), include olapcore_compat.hpp after its own headers,
compile it and link it against libolapcore.so. The modified copy goes to a
temporary folder only; the test is skipped where the reference sources are
absent (the GPU box)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
SRC = os.path.join(REF, "code", "OLAPcore", "test_multi_compute_operator.cpp")
LIBDIR = os.path.join(ROOT, "memory_database_amd")


@pytest.mark.skipif(not os.path.exists(SRC), reason="reference sources absent")
@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_reference_main_links_against_the_library(tmp_path):
    assert os.path.exists(os.path.join(LIBDIR, "libolapcore.so")), "build the library first"
    lines = open(SRC).read().split("\n")
    assert lines[684].startswith("int agg_cwm_dv("), "the reference file moved"
    assert lines[1436].strip() == "}" and lines[1437].startswith("/**")
    assert lines[28].rstrip("\r") == '#include "../include/metadata.h"'
    kept = lines[:29] + ['#include "olapcore_compat.hpp"'] + lines[29:684] + lines[1437:]
    main = tmp_path / "olap_main.cpp"
    main.write_text("\n".join(kept))
    obj = tmp_path / "olap_main.o"
    exe = tmp_path / "olap_main"
    r = subprocess.run(["g++", "-O1", "-std=c++17", "-w", f"-I{ROOT}/include", f"-I{REF}/code",
                        "-c", str(main), "-o", str(obj)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run(["g++", str(obj), "-o", str(exe), f"-L{LIBDIR}", "-lolapcore",
                        f"-Wl,-rpath,{LIBDIR}", "-L/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib",
                        "-lnuma", "-lpthread"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    # every operator the reference's tests call resolves into the library
    syms = subprocess.run(["nm", "-u", str(exe)], capture_output=True, text=True).stdout
    for name in ("OLAPcore_rowwise", "OLAPcore_vwm_dv", "OLAPcore_vwm_sv", "OLAPcore_vwm_dv_numa",
                 "OLAPcore_vwm_sv_numa", "join_cwm_dv", "agg_cwm_dv", "join_cwm_sv", "agg_cwm_sv"):
        assert name in syms, name
