"""The serving loop's host side under ThreadSanitizer: tests/native/
server_race.cpp drives a host-executor server from several threads (submit,
wait with timeouts, cancel, pause, statistics) while its workers run. The
server source is compiled with -fsanitize=thread for the host only; the rest
of the library is linked as built. Skipped when the toolchain or the built
library is absent."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(ROOT, "memory_database_amd")
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC) or not os.path.exists(os.path.join(LIB_DIR, "libolapcore.so")),
                    reason="needs hipcc and the built library")
def test_server_loop_is_race_free_under_tsan(tmp_path):
    exe = tmp_path / "server_race"
    cmd = [HIPCC, "-O1", "-g", "-std=c++17", "-Xarch_host", "-fsanitize=thread",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(LIB_DIR, "csrc"),
           os.path.join(ROOT, "tests", "native", "server_race.cpp"),
           os.path.join(LIB_DIR, "csrc", "olap_server.cpp"),
           "-L", LIB_DIR, "-lolapcore", f"-Wl,-rpath,{LIB_DIR}", "-lpthread", "-o", str(exe)]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr[-3000:]
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 second_deadlock_stack=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300, env=env)
    print(r.stdout)
    assert "ThreadSanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0, (r.stdout, r.stderr[-3000:])
