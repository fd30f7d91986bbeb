"""Host code under AddressSanitizer + UndefinedBehaviorSanitizer (no GPU): the library's OBJ
reader on the reference's Cornell box files and on malformed files, the library's host scene
code (BVH build, wide-tree collapse, packing, objects by id), and the oracle's scene build,
tracers, firefly filter, ray queries and compositors on that scene (oracle/sanitize_check.cpp,
built by `make -C oracle sanitize`). Any sanitizer report aborts the program."""

import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_code_clean_under_asan_ubsan(tmp_path):
    odir = os.path.join(ROOT, "oracle")
    b = subprocess.run(["make", "-s", "-C", odir, "sanitize"], capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stdout + b.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(odir, "_san", "sanitize_check"), os.path.join(ROOT, "tests", "golden", "cornellbox"),
                        str(tmp_path)], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "sanitize_check ok" in r.stdout
