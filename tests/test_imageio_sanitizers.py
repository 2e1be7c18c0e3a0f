"""The image file readers (labeling_amd/csrc/host/imageio.cpp) under
AddressSanitizer and UBSan (tests/cpp/imageio_fuzz_check.cpp): valid PNG, BMP,
TIFF, JPEG and PNM files written by Pillow and by hand, and damaged copies of
each (truncated at many lengths, runs of bytes overwritten with random values),
read as color and as gray: a reader may reject a file, but never touches memory
outside its buffers, and an accepted file gives an image of consistent size."""
import os
import subprocess

import numpy as np
import pytest

PIL = pytest.importorskip("PIL.Image")
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _valid_files(d):
    rng = np.random.default_rng(5)
    rgb = rng.integers(0, 256, (23, 37, 3), dtype=np.uint8)
    gray = rng.integers(0, 256, (19, 21), dtype=np.uint8)
    out = []

    def save(img, name, **kw):
        p = os.path.join(d, name)
        img.save(p, **kw)
        out.append(p)

    im, g = PIL.fromarray(rgb), PIL.fromarray(gray)
    save(im, "a.png")
    save(g, "g.png")
    save(im.convert("P"), "p.png")
    save(im.convert("RGBA"), "rgba.png")
    save(g.convert("1"), "bw.png")
    save(im, "a.bmp")
    save(g, "g.bmp")
    save(im.convert("P"), "p.bmp")
    save(g.convert("1"), "bw.bmp")
    for comp in ("raw", "tiff_lzw", "packbits", "tiff_deflate"):
        save(im, f"a_{comp}.tif", compression=comp)
    save(g, "g_pred.tif", compression="tiff_lzw", tiffinfo={317: 2})
    save(im, "a.jpg", quality=90)
    save(im, "a420.jpg", quality=75, subsampling=2)
    save(g, "g.jpg", quality=85)
    save(im, "a.ppm")
    save(g, "g.pgm")
    return out


def _damaged(d, files):
    rng = np.random.default_rng(9)
    out = []
    for f in files:
        data = open(f, "rb").read()
        base = os.path.basename(f)
        for k, cut in enumerate(sorted(set(int(x) for x in np.linspace(1, len(data) - 1, 12)))):
            p = os.path.join(d, f"t{k}_{base}")
            open(p, "wb").write(data[:cut])
            out.append(p)
        for k in range(24):
            b = bytearray(data)
            at = int(rng.integers(0, len(b)))
            n = int(rng.integers(1, 9))
            for i in range(at, min(len(b), at + n)):
                b[i] = int(rng.integers(0, 256))
            p = os.path.join(d, f"m{k}_{base}")
            open(p, "wb").write(bytes(b))
            out.append(p)
    return out


def test_readers_under_asan_ubsan(tmp_path):
    d = str(tmp_path)
    valid = _valid_files(d)
    files = valid + _damaged(d, valid)
    exe = str(tmp_path / "imageio_fuzz_check")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                        "-I", ROOT, os.path.join(HERE, "cpp", "imageio_fuzz_check.cpp"),
                        os.path.join(ROOT, "labeling_amd", "csrc", "host", "imageio.cpp"), "-o", exe, "-lz"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    r = subprocess.run([exe] + files, capture_output=True, text=True, timeout=600,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0"))
    print(r.stdout[-1000:])
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert r.stdout.strip().endswith("ok")
    accepted = int(r.stdout.split("accepted ")[1].split()[0])
    assert accepted >= 2 * len(valid)   # every valid file reads in both modes
