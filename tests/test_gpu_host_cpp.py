"""The C++ registry mirror (SetLabelingAlg / ILabeling::Label) and the CLI,
driven like the reference's main.cpp, on the GPU: labels equal the oracle's,
error behaviour follows the reference."""
import os
import subprocess

import numpy as np
import pytest

import labeling_amd as L
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(L.LIB_PATH)


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("cpp") / "registry_check")
    cmd = ["g++", "-O1", "-std=c++17", os.path.join(HERE, "cpp", "registry_check.cpp"), "-I",
           os.path.join(PKG, "csrc", "host"), "-L", PKG, "-llabeling_tools", "-llabeling_amd",
           f"-Wl,-rpath,{PKG}", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


# (name, device, expected labels of a binary image): every ALG_LIST entry through
# the factory main.cpp would pick (-g for the GPU classes, none for the CPU ones)
REGISTRY = [
    ("bleq", "gpu", lambda b: O.canonical(O.bfs2d(b, 8)[0])[0], True),
    ("lbeq", "gpu", lambda b: O.canonical(O.bfs2d(b, 4)[0])[0], True),
    ("runeq", "gpu", lambda b: O.runeq2d(b), False),
    ("bin", "gpu", lambda b: O.bin2d(b), False),
    ("bleq", "cpu", lambda b: O.canonical(O.bfs2d(b, 8)[0])[0], True),
    ("he-run", "cpu", lambda b: O.herun2d(b, 4), False),
    ("gr-block", "cpu", lambda b: O.grblock2d(b)[0], False),
    ("ocv", "cpu", lambda b: O.canonical(O.bfs2d(b, 4)[0])[0], False),
    ("runeq", "cpu", lambda b: O.runeq2d(b), False),
]


@pytest.mark.parametrize("entry", REGISTRY, ids=lambda e: f"{e[0]}-{e[1]}")
def test_registry_2d(checker, tmp_path, entry):
    alg, device, expect, canonical = entry
    img = L.bernoulli((301, 517), 0.5, seed=3)
    img[img > 0] = 9      # binarised by RGB2Gray's v > 8 rule
    img[0, :5] = 8        # 8 is background after binarisation
    raw = tmp_path / "in.raw"
    img.tofile(raw)
    out = tmp_path / "out.raw"
    r = subprocess.run([checker, str(raw), "301", "517", str(out), "0", alg, device], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lab = np.fromfile(out, dtype=np.uint32).reshape(img.shape)
    binary = np.where(img > 8, 255, 0).astype(np.uint8)
    got = O.canonical(lab)[0] if canonical else lab
    np.testing.assert_array_equal(got, expect(binary))


def test_registry_3d(checker, tmp_path):
    vol = L.bernoulli((21, 33, 40), 0.3, seed=4)
    raw = tmp_path / "in.raw"
    vol.tofile(raw)
    out = tmp_path / "out.raw"
    r = subprocess.run([checker, str(raw), "33", "40", str(out), "21", "bleq"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lab = np.fromfile(out, dtype=np.uint32).reshape(vol.shape)
    np.testing.assert_array_equal(O.canonical(lab)[0], O.bfs3d(vol, 26)[0])


def _write_pgm(path, img):
    with open(path, "wb") as f:
        f.write(f"P5\n{img.shape[1]} {img.shape[0]}\n255\n".encode())
        f.write(np.ascontiguousarray(img).tobytes())


def test_cli_file_dir_and_3d(tmp_path):
    exe = os.path.join(PKG, "labeling")
    img = L.bernoulli((100, 130), 0.5, seed=5)
    _write_pgm(tmp_path / "a.pgm", img)
    _write_pgm(tmp_path / "b.pgm", img[::-1])
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([exe, "-a", "bleq", "-g", "-i", str(tmp_path / "a.pgm"), "-o", str(out), "-l", "3"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "Min =" in r.stdout, r.stdout
    assert (out / "a.pgm.ppm").exists()
    r = subprocess.run([exe, "-a", "lbeq", "-g", "-c", "8", "-i", str(tmp_path)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0 and "Processing image 2/2" in r.stdout, r.stdout
    sl = tmp_path / "slices"
    sl.mkdir()
    for k in range(5):
        _write_pgm(sl / f"{k:03d}.pgm", L.bernoulli((40, 50), 0.3, seed=6, stream=k))
    r = subprocess.run([exe, "-a", "bleq", "-g", "-3", "-i", str(sl), "-o", str(out)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0 and "Processing time" in r.stdout, r.stdout
    assert (out / "5.ppm").exists()
    r = subprocess.run([exe, "-a", "bleq", "-g", "-c", "4", "-i", str(tmp_path / "a.pgm")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0 and "4-connectivity" in r.stdout
    # he-run has no GPU (OpenCL) factory in the reference: -g fails, the CPU entry runs
    r = subprocess.run([exe, "-a", "he-run", "-g", "-i", str(tmp_path / "a.pgm")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0 and "doesn't support" in r.stdout
    r = subprocess.run([exe, "-a", "he-run", "-i", str(tmp_path / "a.pgm")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0 and "Processing time" in r.stdout, r.stdout
