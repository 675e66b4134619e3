"""The C ABI library loads without a GPU and exports every symbol the
headers in include/ declare (no compute calls here)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from memory_database_amd import olapcore as oc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for h in os.listdir(INCLUDE):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"\btypedef\b[^;]*;", "", text)  # function-pointer types are not symbols
        names |= set(re.findall(r"\b(olap_[a-z0-9_]+)\s*\(", text))
    return names


def exported_symbols(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_library_exports_every_declared_symbol():
    assert os.path.exists(oc.LIB_PATH), "build the library first"
    decl = declared_functions()
    assert len(decl) >= 30
    exported = exported_symbols(oc.LIB_PATH)
    missing = sorted(decl - exported)
    assert not missing, f"declared but not exported: {missing}"


def test_bindings_cover_the_header():
    decl = declared_functions()
    assert decl == set(oc.SIGNATURES), (sorted(decl - set(oc.SIGNATURES)),
                                        sorted(set(oc.SIGNATURES) - decl))


def test_loads_and_fails_cleanly_without_gpu():
    lib = oc.load()
    h = C.c_void_p()
    st = lib.olap_ctx_create(0, C.byref(h))
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        assert st == oc.OK
        lib.olap_ctx_destroy(h)
    else:
        assert st == oc.E_NODEVICE and not h.value


def test_host_only_entry_points():
    import numpy as np

    p = oc.make_plan([0, 0, 16, 0], [0, 0, 1, 0], [0, 0, 0.1, 0])
    assert p.dimvec_nums == 1 and p.orders[0] == 2 and p.group_nums == 16
    with pytest.raises(oc.OlapError):
        oc.make_plan([0, 0, 0, 0], [0, 0, 0, 0], [0, 0, 0, 0])
    with pytest.raises(oc.OlapError):
        oc.make_plan([70000, 70000, 1, 1], [1, 1, 1, 1], [0.1] * 4)  # G > 2^31
    assert oc.legacy_count(np.array([1, 2, 3], np.int64)) == 6
    thr = oc.zipf_table(10, 0.0)
    assert thr[-1] == 2 ** 32
