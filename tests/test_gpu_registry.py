"""GPU parity of the reference registry's other methods (ALG_LIST,
main.cpp:103-130) through the C ABI: bin, runeq, he-run, gr-block, ocv.

Bit-exact against the oracle's restatements (oracle/oracle_registry.c) at
sizes it finishes quickly, and against the closed-form values of
tests/registry_values.py at larger sizes; counts against the BFS labeler.
Shapes cover 1-pixel images and rows, odd sizes, widths around the 64-pixel
words of the run bits, rows of several 4096-pixel chunks (the run-bit kernel
carries a run's state from chunk to chunk), long all-foreground rows (runs
that cross many words) and every density.
"""

import numpy as np
import pytest

import labeling_amd as L
from oracle import pyoracle as O
from tests import registry_values as V

pytestmark = pytest.mark.gpu


def _images():
    out = []
    for (h, w) in [(1, 1), (1, 7), (7, 1), (2, 2), (5, 63), (6, 64), (7, 65), (33, 127), (40, 128), (41, 129),
                   (96, 200), (3, 4097), (9, 5000), (5, 9000), (130, 257)]:
        for p in (0.1, 0.5, 0.9):
            out.append((f"{h}x{w}_p{p}", L.bernoulli((h, w), p, seed=h * 7 + w, stream=int(p * 10))))
    # long runs: full rows crossing many words and chunks, touching rows above in one column only
    img = np.zeros((8, 9000), np.uint8)
    img[1, :] = 255
    img[3, 100:8900] = 255
    img[2, 8899] = 255
    img[5, 4095:4097] = 255
    img[6, 4096] = 255
    img[7, :] = 255
    out.append(("long_runs", img))
    out.append(("blobs_300", L.blobs2d(300, 300, 4.0, 0.5, seed=3)))
    return out


IMAGES = _images()


def _bfs(img, conn):
    lab, n = O.bfs2d(img, conn)
    return O.canonical(lab)[0], n


@pytest.mark.parametrize("case", IMAGES, ids=lambda c: c[0])
def test_bin(labeler, case):
    _, img = case
    labeler.set_mode(L.LABELS_ROOT)
    lab, n, _ = labeler.label2d(img, "bin")
    assert np.array_equal(lab, O.bin2d(img)) and n == int(img.any())


@pytest.mark.parametrize("case", IMAGES, ids=lambda c: c[0])
def test_runeq(labeler, case):
    _, img = case
    h, w = img.shape
    b4, n4 = _bfs(img, 4)
    small = h * w <= 200_000
    for mode, ref in ((L.LABELS_ROOT, False), (L.LABELS_REFERENCE, True)):
        labeler.set_mode(mode)
        lab, n, _ = labeler.label2d(img, "runeq")
        expect = O.runeq2d(img, reference=ref) if small else \
            V.runeq_values(img, (L.ref_pad2(w) // 2) if ref else (w + 1) // 2)
        assert np.array_equal(lab, expect) and n == n4
    labeler.set_mode(L.LABELS_CANONICAL)
    lab, n, _ = labeler.label2d(img, "runeq", 8)
    assert np.array_equal(lab, b4) and n == n4
    labeler.set_mode(L.LABELS_ROOT)


@pytest.mark.parametrize("case", IMAGES, ids=lambda c: c[0])
def test_herun(labeler, case):
    _, img = case
    h, w = img.shape
    for conn in (4, 8):
        b, nb = _bfs(img, conn)
        expect = O.herun2d(img, conn) if h * w <= 70_000 else V.herun_values(img, conn)
        for mode in (L.LABELS_ROOT, L.LABELS_REFERENCE):
            labeler.set_mode(mode)
            lab, n, _ = labeler.label2d(img, "he-run", conn)
            assert np.array_equal(lab, expect) and n == nb, (conn, mode)
        labeler.set_mode(L.LABELS_CANONICAL)
        lab, n, _ = labeler.label2d(img, "he-run", conn)
        assert np.array_equal(lab, b) and n == nb
    labeler.set_mode(L.LABELS_ROOT)
    lab0, _, _ = labeler.label2d(img, "he-run", 0)   # default: 4 (ConPix = coh == COH_8, LabelingAlgs.cpp:140)
    lab4, _, _ = labeler.label2d(img, "he-run", 4)
    assert np.array_equal(lab0, lab4)


@pytest.mark.parametrize("case", IMAGES, ids=lambda c: c[0])
def test_grblock(labeler, case):
    _, img = case
    b8, n8 = _bfs(img, 8)
    labeler.set_mode(L.LABELS_ROOT)
    lab, n, _ = labeler.label2d(img, "gr-block")
    expect = O.grblock2d(img)[0] if img.size <= 200_000 else V.block_order_values(img)
    assert np.array_equal(lab, expect) and n == n8
    labeler.set_mode(L.LABELS_CANONICAL)
    lab, n, _ = labeler.label2d(img, "gr-block")
    assert np.array_equal(lab, b8) and n == n8
    labeler.set_mode(L.LABELS_ROOT)


@pytest.mark.parametrize("case", IMAGES[::3], ids=lambda c: c[0])
def test_ocv(labeler, case):
    """cv::connectedComponents (LabelingAlgs.cpp:39-45): 8 if asked, else 4;
    1..N in raster order of first pixels in every mode."""
    _, img = case
    for conn, want in ((0, 4), (4, 4), (8, 8)):
        b, nb = _bfs(img, want)
        for mode in (L.LABELS_ROOT, L.LABELS_REFERENCE, L.LABELS_CANONICAL):
            labeler.set_mode(mode)
            lab, n, _ = labeler.label2d(img, "ocv", conn)
            assert np.array_equal(lab, b) and n == nb
    labeler.set_mode(L.LABELS_ROOT)


def test_large_2048(labeler):
    """2048 x 2048, p = 0.5: every method against its closed form."""
    img = L.bernoulli((2048, 2048), 0.5, seed=11)
    labeler.set_mode(L.LABELS_ROOT)
    lab, n, _ = labeler.label2d(img, "runeq")
    assert np.array_equal(lab, V.runeq_values(img, 1024))
    lab, n, _ = labeler.label2d(img, "gr-block")
    assert np.array_equal(lab, V.block_order_values(img))
    for conn in (4, 8):
        lab, n, _ = labeler.label2d(img, "he-run", conn)
        assert np.array_equal(lab, V.herun_values(img, conn))


def test_connectivity_and_dimension_errors(labeler):
    img = L.bernoulli((20, 30), 0.5, seed=1)
    for alg in ("runeq", "gr-block"):   # LabelingAlgs.cpp:51,537,1242
        with pytest.raises(L.LabelingError) as e:
            labeler.label2d(img, alg, 4)
        assert e.value.status == L.EUNSUPPORTED
    vol = L.bernoulli((6, 7, 8), 0.5, seed=2)
    for alg in ("runeq", "he-run", "gr-block", "ocv"):   # no 3D entries (main.cpp:103-130)
        with pytest.raises(L.LabelingError) as e:
            labeler.label3d(vol, alg)
        assert e.value.status == L.EUNSUPPORTED
    with pytest.raises(L.LabelingError) as e:
        labeler.label2d(img, "he-run", 6)
    assert e.value.status == L.EINVAL


def test_bin3d(labeler):
    """TOCLBinLabeling3D (LabelingAlgs.cpp:1270-1318): 1 on foreground, cropped or
    in the reference's padded volume (data at +2)."""
    vol = L.bernoulli((9, 17, 33), 0.4, seed=4)
    labeler.set_mode(L.LABELS_ROOT)
    lab, n, _ = labeler.label3d(vol, "bin")
    assert np.array_equal(lab, (vol != 0).astype(np.uint32)) and n == 1
    labeler.set_mode(L.LABELS_REFERENCE_UNCROPPED)
    lab, n, _ = labeler.label3d(vol, "bin", padded=True)
    expect = np.zeros(lab.shape, np.uint32)
    expect[2:2 + 9, 2:2 + 17, 2:2 + 33] = vol != 0
    assert np.array_equal(lab, expect)
    labeler.set_mode(L.LABELS_ROOT)
    lab, n, _ = labeler.label3d(np.zeros((3, 4, 5), np.uint8), "bin")
    assert n == 0 and not lab.any()


def test_device_buffers(labeler):
    """The device-buffer entry point runs the same methods (count on the device)."""
    from tests import hiprt

    img = L.bernoulli((300, 500), 0.45, seed=9)
    d_px = hiprt.to_device(img)
    d_lab = hiprt.DeviceBuffer(300 * 500 * 4)
    d_n = hiprt.DeviceBuffer(4)
    labeler.set_mode(L.LABELS_ROOT)
    for alg, conn, expect in (("runeq", 0, O.runeq2d(img)), ("he-run", 8, O.herun2d(img, 8)),
                              ("gr-block", 0, O.grblock2d(img)[0]), ("bin", 0, O.bin2d(img))):
        labeler.label2d_dev(d_px.addr, 300, 500, 500, d_lab.addr, alg, conn, d_n.addr)
        hiprt.synchronize()
        got = d_lab.download((300, 500), np.uint32)
        assert np.array_equal(got, expect), alg
        n = int(d_n.download((1,), np.uint32)[0])
        assert n == (int(O.canonical(expect)[1]) if alg != "bin" else 1), alg
