"""The C++ host side over the C ABI, without a GPU: include/x3dpt.hpp compiles cleanly with
-Wall -Wextra -pedantic, and the host program x3dcore2_amd/host/render_cornell (built by
x3dcore2_amd/build.py) builds the reference's Cornell box from its OBJ files on the host and
reports the commit's device error as an x3d::error (or commits, where a device is present)."""

import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "x3dcore2_amd", "host", "render_cornell")
OBJ = os.path.join(ROOT, "tests", "golden", "cornellbox")


def test_header_compiles_cleanly(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text('#include "x3dpt.hpp"\nint main() { x3d_render_options o{}; (void)o; return 0; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-pedantic", "-Werror", "-fsyntax-only",
                        "-I" + os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_host_program_builds_scene_on_host(x3d):
    if not os.path.exists(HOST):
        pytest.skip("host program not built (python -m x3dcore2_amd.build)")
    r = subprocess.run([HOST, OBJ, "--host-only"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "no HIP device" in r.stdout or "committed" in r.stdout
    bad = subprocess.run([HOST, OBJ + "_missing", "--host-only"], capture_output=True, text=True, timeout=120)
    assert bad.returncode == 1 and "cannot open" in bad.stderr
