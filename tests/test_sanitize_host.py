"""Host C++ of the library (builder, topology, workload generator, minima
database) built with g++ -fsanitize=address,undefined and run over the typed
fixtures, 128 generated workload systems and a database sequence: no
sanitizer report, exit status 0.  Skipped where g++ cannot link the
sanitizer runtimes."""
import glob
import os
import shutil
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "firecore_amd", "csrc")
TYPED = os.path.join(HERE, "golden", "typed")
SOURCES = ["builder.cpp", "database.cpp", "topology.cpp", "molgen.cpp"]


def test_host_code_clean_under_sanitizers():
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "sanitize_host")
        cmd = [gxx, "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
               "-fno-sanitize-recover=undefined", "-I" + CSRC, os.path.join(HERE, "sanitize_host.cpp")]
        cmd += [os.path.join(CSRC, s) for s in SOURCES] + ["-o", exe, "-lpthread"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            if "asan" in r.stderr or "ubsan" in r.stderr:
                pytest.skip("sanitizer runtimes unavailable: " + r.stderr[-300:])
            raise AssertionError(r.stderr[-3000:])
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
                   UBSAN_OPTIONS="print_stacktrace=1")
        xyz = sorted(glob.glob(os.path.join(TYPED, "*.xyz")))
        r = subprocess.run([exe, TYPED] + xyz, capture_output=True, text=True, env=env, timeout=300)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
        assert "ERROR" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
        assert "builder: %d built" % len(xyz) in r.stdout, r.stdout
        assert "database:" in r.stdout and "gen fail" not in r.stdout, r.stdout
