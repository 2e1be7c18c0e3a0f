"""Model check of the concurrent unions on the CPU (tests/cpp/rem_interleave_check.cpp):
the tile kernels' LDS loop (Rem's splicing, lds_rem_step) and the global
merges' unite (finds by cacheable loads, root check by a device-coherent load,
root link by atomicMin) keep every union under random interleavings of single
loads and atomicMins: the LDS loop with sequentially consistent memory, the
global unite also with loads that return any earlier value of their location.
The 2D loop's start of a union from the node where the lane's previous union
met (lds-top) keeps every union too.  The LDS loop's former last step (both
original nodes pointed at the node where the sets met) loses a union on a
recorded schedule of 12 lanes, which the check replays: it must fail there
while the loop without that step keeps it."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_union_interleavings(tmp_path):
    exe = str(tmp_path / "rem_check")
    r = subprocess.run(["g++", "-std=c++17", "-O2", os.path.join(HERE, "cpp", "rem_interleave_check.cpp"), "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe, "60000"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "device variants exact" in r.stdout
