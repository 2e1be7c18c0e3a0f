"""The oracle's C code (oracle/oracle.c, oracle/oracle_registry.c) under
AddressSanitizer and UBSan: every entry point over random images and volumes
of small and odd shapes (tests/cpp/oracle_sanitize_check.c), with no memory
error or undefined behaviour and the exact labelers equal to the BFS labeler."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_oracle_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "oracle_sanitize_check")
    src = [os.path.join(HERE, "cpp", "oracle_sanitize_check.c"), os.path.join(ROOT, "oracle", "oracle.c"),
           os.path.join(ROOT, "oracle", "oracle_registry.c")]
    r = subprocess.run(["gcc", "-std=c11", "-O1", "-g", "-fopenmp", "-fsanitize=address,undefined",
                        "-fno-sanitize-recover=all", "-I", ROOT] + src + ["-o", exe, "-lm"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", OMP_NUM_THREADS="2")
    r = subprocess.run([exe, "300"], capture_output=True, text=True, timeout=600, env=env)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "ok (0 failures)" in r.stdout
