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


def _color_input(h, w, seed):
    """An RGB image whose gray levels straddle the v > 8 threshold."""
    rng = np.random.default_rng(seed)
    rgb = rng.integers(0, 14, (h, w, 3), dtype=np.uint8)
    rgb[rng.random((h, w)) < 0.3] += 100
    return rgb


def _binary_2d(rgb):
    """The reference's 2D input: cv::imread BGR through cvtColor(COLOR_RGB2GRAY)
    (weights on B, G, R), then v > 8 (LabelingTools.cpp:39-49)."""
    b, g, r = (rgb[..., 2 - k].astype(np.int64) for k in range(3))
    gray = (b * 4899 + g * 9617 + r * 1868 + 8192) >> 14
    return (gray > 8).astype(np.uint8)


def _color_ids(rgb):
    """Canonical labels of a colourised label image (a colour per component)."""
    flat = rgb.reshape(-1, 3).astype(np.int64)
    key = (flat[:, 0] << 16) | (flat[:, 1] << 8) | flat[:, 2]
    ids = np.unique(key, return_inverse=True)[1].reshape(rgb.shape[:2]) + 1
    ids[key.reshape(rgb.shape[:2]) == 0] = 0
    return O.canonical(ids.astype(np.uint32))[0]


def test_cli_formats_dir_and_3d(tmp_path):
    """The driver's file handling (main.cpp:134-152,172-212,250-340,466-490):
    every format the reference's directory scan accepts goes in, the labels come
    out colourised under the input's name and format; 3D slices come out as
    1.png, 2.png, ...  Checked: the colour partition equals the oracle's
    components of the reference's binarisation of the decoded input."""
    from PIL import Image

    exe = os.path.join(PKG, "labeling")
    rgb = _color_input(100, 130, 5)
    inputs = {"a.png": rgb, "b.bmp": rgb[::-1], "c.tif": rgb[:, ::-1], "d.tiff": rgb[::2], "e.ppm": rgb[:50]}
    for name, a in inputs.items():
        Image.fromarray(np.ascontiguousarray(a)).save(tmp_path / name)
    Image.fromarray(np.ascontiguousarray(rgb)).save(tmp_path / "f.jpg", quality=95)
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([exe, "-a", "bleq", "-g", "-i", str(tmp_path / "a.png"), "-o", str(out), "-l", "3"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "Min =" in r.stdout, r.stdout
    lab = np.asarray(Image.open(out / "a.png").convert("RGB"))
    np.testing.assert_array_equal(_color_ids(lab), O.canonical(O.bfs2d(_binary_2d(rgb), 8)[0])[0])
    r = subprocess.run([exe, "-a", "lbeq", "-g", "-c", "8", "-i", str(tmp_path), "-o", str(out)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0 and "Processing image 6/6" in r.stdout, r.stdout
    for name, a in inputs.items():   # lossless formats: exact partition
        lab = np.asarray(Image.open(out / name).convert("RGB"))
        np.testing.assert_array_equal(_color_ids(lab), O.canonical(O.bfs2d(_binary_2d(a), 8)[0])[0])
    assert Image.open(out / "f.jpg").size == (130, 100)
    sl = tmp_path / "slices"
    sl.mkdir()
    vol = np.stack([L.bernoulli((40, 50), 0.3, seed=6, stream=k) for k in range(5)], axis=-1)   # {rows, cols, slices}
    for k in range(5):
        Image.fromarray(vol[..., k]).save(sl / f"{k:03d}.png")
    out3 = tmp_path / "out3"
    out3.mkdir()
    r = subprocess.run([exe, "-a", "bleq", "-g", "-3", "-i", str(sl), "-o", str(out3)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0 and "Processing time" in r.stdout, r.stdout
    planes = np.stack([np.asarray(Image.open(out3 / f"{k + 1}.png").convert("RGB")) for k in range(5)], axis=2)
    flat = planes.reshape(40, 50 * 5, 3)
    ids = _color_ids(flat).reshape(40, 50, 5)
    np.testing.assert_array_equal(O.canonical(ids.astype(np.uint32))[0], O.canonical(O.bfs3d(vol, 26)[0])[0])
    r = subprocess.run([exe, "-a", "bleq", "-g", "-c", "4", "-i", str(tmp_path / "a.png")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0 and "4-connectivity" in r.stdout
    # he-run has no GPU (OpenCL) factory in the reference: -g fails, the CPU entry runs
    r = subprocess.run([exe, "-a", "he-run", "-g", "-i", str(tmp_path / "a.png")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0 and "doesn't support" in r.stdout
    r = subprocess.run([exe, "-a", "he-run", "-i", str(tmp_path / "a.png")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0 and "Processing time" in r.stdout, r.stdout
