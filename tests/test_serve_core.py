"""The server's scheduler core (labeling_amd/csrc/serve_core.hpp) on the CPU
with a simulated device: completion statuses under injected faults, launch and
completion-query errors, admission by arena footprint, fusing, and arena
accounting -- built with AddressSanitizer + UBSan, and with ThreadSanitizer."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "cpp", "serve_core_check.cpp")


@pytest.mark.parametrize("san", ["address,undefined", "thread"])
def test_serve_core(tmp_path, san):
    exe = str(tmp_path / "serve_core_check")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-pthread", f"-fsanitize={san}", "-fno-sanitize-recover=all",
           "-I", ROOT, SRC, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ok (0 failures)" in r.stdout
