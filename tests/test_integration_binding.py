"""The SAFIRE-side binding (tests/integration/hip_binding.hpp: the adapter a SAFIRE build adds
to call this library) compiles against include/afq.h, links with libafq_hip.so, and maps a walker buffer in the reference's
row layout onto afq_walkers (walker strides = the row, each property at its data_displ
column).  Host only: the program never calls into the GPU library."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "integration", "compile_check.cpp")
LIBDIR = os.path.join(ROOT, "safire_amd")


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
@pytest.mark.skipif(not os.path.exists(os.path.join(LIBDIR, "libafq_hip.so")),
                    reason="library not built")
def test_binding_compiles_links_and_maps_the_walker_buffer(tmp_path):
    exe = tmp_path / "binding_check"
    cmd = ["g++", "-std=c++20", "-Wall", "-Wextra", "-Werror", "-Wno-unused-parameter",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "tests", "integration"),
           SRC, "-o", str(exe), "-L", LIBDIR, "-lafq_hip", "-Wl,-rpath," + LIBDIR,
           "-Wl,--allow-shlib-undefined"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    # the program maps a host-side walker buffer; it never calls into the library
    r = subprocess.run([str(exe)], capture_output=True, text=True,
                       env={**os.environ, "HIP_VISIBLE_DEVICES": ""})
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ok" in r.stdout
