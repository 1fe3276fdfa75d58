"""The static library links into a plain C++ program with the HIP runtime
(the link a Rust crate's build.rs makes): host entry points give the same
results as the shared library's, and context creation without a GPU fails
with the documented error instead of crashing."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT

STATIC = os.path.join(ROOT, "jixel_amd", "libjxg.a")


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.exists(STATIC),
                    reason="needs g++ and a built libjxg.a")
def test_static_library_links_and_runs(tmp_path):
    import jixel_amd as J
    exe = tmp_path / "static_link"
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "native", "static_link.cpp"), STATIC,
                           "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib",
                           "-o", str(exe)])
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    words = out.stdout.split()
    scale, table0, ctx_rc = float(words[1]), float(words[3]), int(words[5])
    p = J.frame_params(J.EncodeConfig().with_distance(2.35))
    assert abs(scale - p.scale) <= 1e-6 * p.scale
    assert abs(table0 - float(J.default_quant_tables(2.35)[0])) <= 1e-6
    assert ctx_rc in (0, -7)  # a gfx950 device, or none
