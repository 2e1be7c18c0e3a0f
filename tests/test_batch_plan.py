"""The batch planner (labeling_amd/csrc/batch_plan.hpp) on the CPU, under
AddressSanitizer + UBSan, for the 4096-image batched workload plus degenerate
shapes, every tile shape (the autotuner's re-plan sequence) and both canonical
settings: every tile / final-workgroup / scan-chunk id a batched kernel grid
launches maps inside its image's grid, tables and workspace slices fit the
buffers sized for the largest plan (tests/cpp/batch_plan_check.cpp)."""
import os
import subprocess

import labeling_amd as L

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_batch_plan(tmp_path):
    shapes = tmp_path / "shapes.txt"
    extra = [(1, 1), (1, 3), (3, 1), (2047, 1), (1, 2047), (5, 7), (2, 65535)]
    shapes.write_text("".join(f"{h} {w}\n" for h, w, _ in L.batched_workload(4096, 0)) +
                      "".join(f"{h} {w}\n" for h, w in extra))
    exe = str(tmp_path / "batch_plan_check")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                        "-I", ROOT, os.path.join(HERE, "cpp", "batch_plan_check.cpp"), "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe, str(shapes)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "ok (0 failures)" in r.stdout
