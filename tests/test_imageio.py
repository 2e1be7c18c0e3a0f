"""The CLI's image files (labeling_amd/csrc/host/imageio.cpp) on the CPU, against
Pillow as an independent decoder / encoder:

* lossless formats (PNG of every color type and bit depth, interlaced or not;
  BMP 1/4/8/16/24/32-bit and RLE8; TIFF raw / LZW / PackBits / Deflate, with
  and without the horizontal predictor; PNM P1-P6): decoded pixels equal
  Pillow's exactly, after cv::imread's conversions (BGR, high byte of 16-bit
  samples, gray weights of imageio.hpp);
* JPEG (baseline; 4:4:4, 4:2:2, 4:2:0 and gray, restart intervals): within one
  level of Pillow's IJG decoder on at least 99% of samples (the inverse DCT is
  computed in floating point, IJG's in fixed point), never off by more than 3;
* writers: PNG / BMP / TIFF / PNM read back by Pillow bit-exactly, JPEG within
  the loss of quality 95.
"""
import ctypes
import os
import struct
import zlib

import numpy as np
import pytest

import labeling_amd as L

PIL = pytest.importorskip("PIL.Image")
TOOLS = os.path.join(os.path.dirname(L.LIB_PATH), "liblabeling_tools.so")


@pytest.fixture(scope="module")
def lib():
    L.load()   # the tools library links the labeling library
    t = ctypes.CDLL(TOOLS)
    t.lt_image_read.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.c_void_p,
                                ctypes.c_size_t]
    t.lt_image_write.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    t.lt_image_error.restype = ctypes.c_char_p
    return t


def read(lib, path, gray=False):
    h, w, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    assert lib.lt_image_read(str(path).encode(), int(gray), ctypes.byref(h), ctypes.byref(w), ctypes.byref(c),
                             None, 0) == 0, lib.lt_image_error()
    out = np.empty((h.value, w.value, c.value), np.uint8)
    assert lib.lt_image_read(str(path).encode(), int(gray), ctypes.byref(h), ctypes.byref(w), ctypes.byref(c),
                             out.ctypes.data, out.nbytes) == 0
    return out


def read_fails(lib, path, gray=False):
    h, w, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    return lib.lt_image_read(str(path).encode(), int(gray), ctypes.byref(h), ctypes.byref(w), ctypes.byref(c),
                             None, 0) != 0


def rng_img(shape, seed):
    return np.random.default_rng(seed).integers(0, 256, shape, dtype=np.uint8)


def bgr_of(pil_img):
    rgb = np.asarray(pil_img.convert("RGB"))
    return rgb[..., ::-1]


def gray14(rgb):
    r, g, b = (rgb[..., k].astype(np.int64) for k in range(3))
    return ((r * 4899 + g * 9617 + b * 1868 + 8192) >> 14).astype(np.uint8)


def gray15(rgb):
    r, g, b = (rgb[..., k].astype(np.int64) for k in range(3))
    return ((r * 9798 + g * 19235 + b * 3735 + 16384) >> 15).astype(np.uint8)


# ---- PNG ---------------------------------------------------------------------------

def write_png(path, samples, ctype, depth, interlace=False, palette=None):
    """A PNG writer for the cases Pillow cannot write: Adam7 interlacing and
    1/2/4-bit gray.  samples: (h, w, channels) integers."""
    h, w = samples.shape[:2]
    ch = samples.shape[2]
    passes = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)] \
        if interlace else [(0, 0, 1, 1)]
    raw = bytearray()
    for x0, y0, dx, dy in passes:
        sub = samples[y0::dy, x0::dx]
        if sub.size == 0:
            continue
        for row in sub:
            raw.append(0)
            if depth < 8:
                bits = "".join(format(int(v), f"0{depth}b") for v in row.reshape(-1))
                bits += "0" * (-len(bits) % 8)
                raw += int(bits, 2).to_bytes(len(bits) // 8, "big") if bits else b""
            elif depth == 8:
                raw += row.astype(np.uint8).tobytes()
            else:
                raw += row.astype(">u2").tobytes()

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)

    data = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, int(interlace)))
    if palette is not None:
        data += chunk(b"PLTE", palette.astype(np.uint8).tobytes())
    data += chunk(b"IDAT", zlib.compress(bytes(raw))) + chunk(b"IEND", b"")
    open(path, "wb").write(data)


@pytest.mark.parametrize("mode", ["L", "LA", "RGB", "RGBA", "P", "1", "I;16"])
def test_png_pillow_modes(lib, tmp_path, mode):
    base = rng_img((37, 53, 4), 1)
    if mode in ("L", "1"):
        im = PIL.fromarray(base[..., 0]).convert(mode)
    elif mode == "LA":
        im = PIL.fromarray(base[..., :2], "LA")
    elif mode == "RGB":
        im = PIL.fromarray(base[..., :3])
    elif mode == "RGBA":
        im = PIL.fromarray(base, "RGBA")
    elif mode == "P":
        im = PIL.fromarray(base[..., :3]).quantize(colors=37)
    else:
        im = PIL.fromarray((base[..., 0].astype(np.uint16) * 257 + base[..., 1]).astype(np.uint16))
    p = tmp_path / "a.png"
    im.save(p)
    got = read(lib, p)
    if mode == "I;16":
        v8 = (np.asarray(im).astype(np.uint16) >> 8).astype(np.uint8)
        expect = np.repeat(v8[..., None], 3, axis=2)
    else:
        expect = bgr_of(im)
    assert np.array_equal(got, expect)
    g = read(lib, p, gray=True)[..., 0]
    rgb = expect[..., ::-1]
    gray_src = mode in ("L", "LA", "1", "I;16")
    assert np.array_equal(g, rgb[..., 0] if gray_src else gray15(rgb))


@pytest.mark.parametrize("depth", [1, 2, 4, 8, 16])
@pytest.mark.parametrize("interlace", [False, True])
def test_png_gray_depths_and_adam7(lib, tmp_path, depth, interlace):
    for h, w in ((1, 1), (5, 9), (17, 31)):
        s = np.random.default_rng(depth * 100 + h).integers(0, 1 << depth, (h, w, 1))
        p = tmp_path / f"g{depth}_{h}.png"
        write_png(p, s, 0, depth, interlace)
        got = read(lib, p)[..., 0]
        want = (s[..., 0] * 255 // ((1 << depth) - 1)) if depth < 8 else (s[..., 0] >> (depth - 8))
        assert np.array_equal(got, want.astype(np.uint8)), (depth, interlace, h, w)
        if depth < 16:   # Pillow's own decode agrees
            assert np.array_equal(np.asarray(PIL.open(p).convert("L")), want.astype(np.uint8))


@pytest.mark.parametrize("interlace", [False, True])
def test_png_rgb16_palette_adam7(lib, tmp_path, interlace):
    s = np.random.default_rng(3).integers(0, 65536, (19, 23, 3))
    p = tmp_path / "c16.png"
    write_png(p, s, 2, 16, interlace)
    assert np.array_equal(read(lib, p), (s >> 8).astype(np.uint8)[..., ::-1])
    pal = np.random.default_rng(4).integers(0, 256, (16, 3))
    idx = np.random.default_rng(5).integers(0, 16, (19, 23, 1))
    p = tmp_path / "p4.png"
    write_png(p, idx, 3, 4, interlace, palette=pal)
    assert np.array_equal(read(lib, p), pal[idx[..., 0]].astype(np.uint8)[..., ::-1])
    assert np.array_equal(bgr_of(PIL.open(p)), pal[idx[..., 0]].astype(np.uint8)[..., ::-1])


# ---- BMP ---------------------------------------------------------------------------

@pytest.mark.parametrize("mode", ["1", "L", "P", "RGB", "RGBA"])
def test_bmp(lib, tmp_path, mode):
    base = rng_img((29, 41, 4), 7)
    im = {"1": lambda: PIL.fromarray(base[..., 0]).convert("1"), "L": lambda: PIL.fromarray(base[..., 0]),
          "P": lambda: PIL.fromarray(base[..., :3]).quantize(colors=200),
          "RGB": lambda: PIL.fromarray(base[..., :3]), "RGBA": lambda: PIL.fromarray(base, "RGBA")}[mode]()
    p = tmp_path / "a.bmp"
    im.save(p)
    assert np.array_equal(read(lib, p), bgr_of(PIL.open(p)))
    assert np.array_equal(read(lib, p, gray=True)[..., 0], gray14(np.asarray(PIL.open(p).convert("RGB"))))


def test_bmp_rle8_and_16bit(lib, tmp_path):
    # RLE8, 3 x 4, bottom-up: rows 2, 1, 0
    pal = bytes([0, 0, 0, 0, 255, 0, 0, 0, 0, 255, 0, 0, 0, 0, 255, 0])   # BGR0: black, blue, green, red
    rle = bytes([4, 1, 0, 0,   0, 4, 2, 3, 3, 2, 0, 0,   2, 3, 2, 0, 0, 1])
    hdr = struct.pack("<IiiHHIIiiII", 40, 4, 3, 1, 8, 1, len(rle), 0, 0, 4, 0)
    off = 14 + 40 + len(pal)
    data = b"BM" + struct.pack("<IHHI", off + len(rle), 0, 0, off) + hdr + pal + rle
    p = tmp_path / "r.bmp"
    p.write_bytes(data)
    colors = np.array([[0, 0, 0], [255, 0, 0], [0, 255, 0], [0, 0, 255]], np.uint8)   # BGR
    want = np.array([[3, 3, 0, 0], [2, 3, 3, 2], [1, 1, 1, 1]])
    assert np.array_equal(read(lib, p), colors[want])
    assert np.array_equal(bgr_of(PIL.open(p)), colors[want])
    # 16-bit 5-5-5, top-down
    v = np.random.default_rng(8).integers(0, 1 << 15, (5, 6)).astype(np.uint16)
    hdr = struct.pack("<IiiHHIIiiII", 40, 6, -5, 1, 16, 0, 0, 0, 0, 0, 0)
    rows = b"".join(r.astype("<u2").tobytes() + b"\0" * 0 for r in v)   # 12 bytes per row: already 4-aligned
    data = b"BM" + struct.pack("<IHHI", 54 + len(rows), 0, 0, 54) + hdr + rows
    p = tmp_path / "s.bmp"
    p.write_bytes(data)
    got = read(lib, p)
    want = np.stack([((v >> s) & 31).astype(np.int64) * 255 // 31 for s in (0, 5, 10)], axis=-1)   # B, G, R
    assert np.abs(got.astype(np.int64) - want).max() <= 1


# ---- TIFF --------------------------------------------------------------------------

@pytest.mark.parametrize("comp", ["raw", "tiff_lzw", "packbits", "tiff_deflate"])
@pytest.mark.parametrize("mode", ["1", "L", "RGB", "RGBA", "P", "I;16"])
def test_tiff(lib, tmp_path, comp, mode):
    base = rng_img((31, 47, 4), 9)
    base[5:20, 10:30] = 77   # runs for the run-length and dictionary coders
    if mode == "I;16":
        im = PIL.fromarray((base[..., 0].astype(np.uint16) * 256 + base[..., 1]).astype(np.uint16))
    elif mode == "P":
        im = PIL.fromarray(base[..., :3]).quantize(colors=100)
    elif mode == "RGBA":
        im = PIL.fromarray(base, "RGBA")
    else:
        im = PIL.fromarray(base[..., :3] if mode == "RGB" else base[..., 0]).convert(mode)
    p = tmp_path / "a.tif"
    try:
        im.save(p, compression=None if comp == "raw" else comp)
    except (OSError, ValueError) as e:
        pytest.skip(f"Pillow cannot write {mode} with {comp}: {e}")
    got = read(lib, p)
    if mode == "I;16":
        v8 = (np.asarray(PIL.open(p)).astype(np.uint16) >> 8).astype(np.uint8)
        assert np.array_equal(got, np.repeat(v8[..., None], 3, axis=2))
    else:
        assert np.array_equal(got, bgr_of(PIL.open(p)))


def test_tiff_predictor(lib, tmp_path):
    im = PIL.fromarray(rng_img((24, 40, 3), 10))
    p = tmp_path / "pred.tif"
    im.save(p, compression="tiff_lzw", tiffinfo={317: 2})
    if PIL.open(p).tag_v2.get(317) != 2:
        pytest.skip("Pillow did not write the predictor")
    assert np.array_equal(read(lib, p), bgr_of(PIL.open(p)))


# ---- PNM ---------------------------------------------------------------------------

def test_pnm(lib, tmp_path):
    g = rng_img((13, 17), 11)
    c = rng_img((13, 17, 3), 12)
    PIL.fromarray(g).save(tmp_path / "a.pgm")
    PIL.fromarray(c).save(tmp_path / "a.ppm")
    assert np.array_equal(read(lib, tmp_path / "a.pgm")[..., 0], g)
    assert np.array_equal(read(lib, tmp_path / "a.ppm"), c[..., ::-1])
    # plain (ASCII) forms, bitmaps (1 = black) and 16-bit samples
    (tmp_path / "p2.pgm").write_text("P2\n# c\n3 2\n255\n0 10 255\n7 8 9\n")
    assert read(lib, tmp_path / "p2.pgm")[..., 0].tolist() == [[0, 10, 255], [7, 8, 9]]
    (tmp_path / "p3.ppm").write_text("P3 2 1 255 1 2 3 4 5 6\n")
    assert read(lib, tmp_path / "p3.ppm").tolist() == [[[3, 2, 1], [6, 5, 4]]]
    (tmp_path / "p1.pbm").write_text("P1\n3 2\n1 0 1\n0 1 0\n")
    assert read(lib, tmp_path / "p1.pbm")[..., 0].tolist() == [[0, 255, 0], [255, 0, 255]]
    (tmp_path / "p4.pbm").write_bytes(b"P4\n3 2\n" + bytes([0b10100000, 0b01000000]))
    assert read(lib, tmp_path / "p4.pbm")[..., 0].tolist() == [[0, 255, 0], [255, 0, 255]]
    v = np.array([[0, 1000, 65535]], np.uint16)
    (tmp_path / "p16.pgm").write_bytes(b"P5 3 1 65535\n" + v.astype(">u2").tobytes())
    assert read(lib, tmp_path / "p16.pgm")[..., 0].tolist() == [[0, 1000 * 255 // 65535, 255]]


# ---- JPEG --------------------------------------------------------------------------

def smooth(shape, seed):
    """A smooth color image (JPEG's intended content) with some texture."""
    rng = np.random.default_rng(seed)
    h, w = shape
    y, x = np.mgrid[0:h, 0:w]
    out = np.stack([128 + 100 * np.sin(x / (7 + 3 * k) + y / (11 + k) + k) for k in range(3)], -1)
    out += rng.normal(0, 6, out.shape)
    return np.clip(out, 0, 255).astype(np.uint8)


@pytest.mark.parametrize("sub", [0, 1, 2])   # Pillow: 0 = 4:4:4, 1 = 4:2:2, 2 = 4:2:0
@pytest.mark.parametrize("shape", [(64, 64), (37, 53), (8, 8), (1, 17), (123, 9)])
def test_jpeg_decode(lib, tmp_path, sub, shape):
    src = smooth(shape, shape[0] * 10 + sub)
    p = tmp_path / "a.jpg"
    PIL.fromarray(src).save(p, quality=90, subsampling=sub)
    got = read(lib, p).astype(np.int64)
    want = bgr_of(PIL.open(p)).astype(np.int64)
    d = np.abs(got - want)
    assert d.max() <= 3 and (d <= 1).mean() >= 0.99, (d.max(), (d <= 1).mean())
    # gray reads take the decoded luma channel (IJG's grayscale output)
    yl = np.asarray(PIL.open(p).convert("YCbCr"))[..., 0].astype(np.int64) if sub == 0 else None
    g = read(lib, p, gray=True)[..., 0].astype(np.int64)
    if yl is not None:
        assert np.abs(g - yl).max() <= 3


def test_jpeg_gray_restart_and_progressive(lib, tmp_path):
    src = smooth((50, 70), 5)[..., 0]
    p = tmp_path / "g.jpg"
    PIL.fromarray(src).save(p, quality=85)
    d = np.abs(read(lib, p)[..., 0].astype(np.int64) - np.asarray(PIL.open(p)).astype(np.int64))
    assert d.max() <= 3 and (d <= 1).mean() >= 0.99
    # restart markers (DRI), written by Pillow's libjpeg when asked
    p = tmp_path / "r.jpg"
    PIL.fromarray(smooth((48, 80), 6)).save(p, quality=90, restart_marker_blocks=3)
    d = np.abs(read(lib, p).astype(np.int64) - bgr_of(PIL.open(p)).astype(np.int64))
    assert d.max() <= 3 and (d <= 1).mean() >= 0.99
    p = tmp_path / "prog.jpg"
    PIL.fromarray(smooth((32, 32), 7)).save(p, progressive=True)
    assert read_fails(lib, p)


# ---- writers -----------------------------------------------------------------------

@pytest.mark.parametrize("ext", [".png", ".bmp", ".tif", ".ppm", ".pgm"])
def test_writers_lossless(lib, tmp_path, ext):
    bgr = rng_img((21, 34, 3), 13)
    p = tmp_path / f"w{ext}"
    assert lib.lt_image_write(str(p).encode(), bgr.ctypes.data, 21, 34, 3) == 0, lib.lt_image_error()
    back = PIL.open(p)
    if ext == ".pgm":
        assert np.array_equal(np.asarray(back), gray14(bgr[..., ::-1]))
    else:
        assert np.array_equal(bgr_of(back), bgr)
    assert np.array_equal(read(lib, p), bgr if ext != ".pgm" else np.repeat(gray14(bgr[..., ::-1])[..., None], 3, 2))


def test_jpeg_writer(lib, tmp_path):
    bgr = smooth((45, 61), 14)
    p = tmp_path / "w.jpg"
    assert lib.lt_image_write(str(p).encode(), bgr.ctypes.data, 45, 61, 3) == 0, lib.lt_image_error()
    back = bgr_of(PIL.open(p)).astype(np.float64)
    psnr = 10 * np.log10(255 ** 2 / np.mean((back - bgr) ** 2))
    assert psnr > 32, psnr
    d = np.abs(read(lib, p).astype(np.int64) - back.astype(np.int64))
    assert d.max() <= 3
    g = smooth((16, 24), 15)[..., :1].copy()
    p = tmp_path / "g.jpg"
    assert lib.lt_image_write(str(p).encode(), g.ctypes.data, 16, 24, 1) == 0
    assert PIL.open(p).mode == "L"


def test_unreadable(lib, tmp_path):
    (tmp_path / "x.png").write_bytes(b"\x89PNG\r\n\x1a\nbroken")
    assert read_fails(lib, tmp_path / "x.png")
    (tmp_path / "y.bin").write_bytes(b"hello")
    assert read_fails(lib, tmp_path / "y.bin")
    assert read_fails(lib, tmp_path / "missing.png")
