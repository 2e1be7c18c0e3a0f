"""Batched 2D block equivalence: many device images of uneven sizes labelled by
one hipGraph launch must give, for every image, exactly the labels and count of
the single-image entry point (and of the oracle)."""
import ctypes

import numpy as np
import pytest

import labeling_amd as L
from oracle import pyoracle as O
from tests import hiprt

pytestmark = pytest.mark.gpu


ImageDesc = L.ImageDesc   # the package's lbl_image2d (its argtypes are declared on load)


def _images(n, seed, max_side=300):
    out = []
    for i in range(n):
        h = 1 + L.synth_u64(seed, 7, 3 * i) % max_side
        w = 1 + L.synth_u64(seed, 7, 3 * i + 1) % max_side
        p = 0.05 + 0.9 * (L.synth_u64(seed, 7, 3 * i + 2) % 1000) / 1000
        out.append(L.bernoulli((h, w), p, seed=seed, stream=i))
    return out


def _upload(imgs, pitch_pad=0):
    bufs, descs = [], (ImageDesc * len(imgs))()
    for i, img in enumerate(imgs):
        h, w = img.shape
        pitch = w + pitch_pad
        host = np.zeros((h, pitch), np.uint8)
        host[:, :w] = img
        d_px = hiprt.to_device(host)
        d_lab = hiprt.DeviceBuffer(h * w * 4)
        bufs.append((d_px, d_lab))
        descs[i] = ImageDesc(d_px.addr, d_lab.addr, h, w, pitch)
    return bufs, descs


@pytest.mark.parametrize("flags", [0, 1, 2], ids=["graph", "autotune", "eager"])
def test_batch_matches_single_image(labeler, flags):
    lib = labeler.lib
    imgs = _images(257, seed=11)
    bufs, descs = _upload(imgs, pitch_pad=flags)   # pitch 16-aligned or not
    d_n = hiprt.DeviceBuffer(4 * len(imgs))
    b = ctypes.c_void_p()
    st = lib.lbl_batch_create(labeler.handle, L.BLEQ, descs, len(imgs), 0, d_n.addr, flags, ctypes.byref(b))
    assert st == L.OK, labeler.lib.lbl_last_error(labeler.handle)
    try:
        tile, tw, th, fl = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_int()
        assert lib.lbl_batch_config(b, ctypes.byref(tile), ctypes.byref(tw), ctypes.byref(th), ctypes.byref(fl)) == 0
        print(f"[batch flags={flags}] tile {tile.value} ({tw.value}x{th.value} blocks) flatten={fl.value}", flush=True)
        for rep in range(3):   # replays with eager work and syncs between them
            for _, d_lab in bufs:
                d_lab.zero()
            assert lib.lbl_batch_run(b, None) == L.OK
            hiprt.synchronize()
            counts = d_n.download((len(imgs),), np.uint32)
            for i, (img, (_, d_lab)) in enumerate(zip(imgs, bufs)):
                got = d_lab.download(img.shape, np.uint32)
                single, n, _ = labeler.label2d(img, "bleq")
                np.testing.assert_array_equal(got, single, err_msg=f"image {i} {img.shape} replay {rep}")
                assert counts[i] == n
                if rep == 0 and i % 16 == 0:
                    np.testing.assert_array_equal(O.canonical(got)[0], O.bfs2d(img, 8)[0])
    finally:
        lib.lbl_batch_destroy(b)


def test_batch_rejects_lbeq_and_bad_images(labeler):
    lib = labeler.lib
    imgs = _images(3, seed=2)
    bufs, descs = _upload(imgs)
    b = ctypes.c_void_p()
    assert lib.lbl_batch_create(labeler.handle, L.LBEQ, descs, 3, 0, None, 0, ctypes.byref(b)) == L.EUNSUPPORTED
    descs[1].h = 0
    assert lib.lbl_batch_create(labeler.handle, L.BLEQ, descs, 3, 0, None, 0, ctypes.byref(b)) == L.EINVAL
    assert lib.lbl_batch_create(labeler.handle, L.BLEQ, descs, 0, 0, None, 0, ctypes.byref(b)) == L.EINVAL


def test_batch_of_mixed_large_images(labeler):
    """A few images of the batched workload's size range (64..2048 per side)."""
    lib = labeler.lib
    sides = [(64, 2048), (2048, 64), (1999, 1031), (517, 733), (2048, 2048), (64, 64), (1000, 1999)]
    imgs = [L.bernoulli(s, 0.05 + 0.15 * k, seed=3, stream=k) for k, s in enumerate(sides)]
    bufs, descs = _upload(imgs)
    d_n = hiprt.DeviceBuffer(4 * len(imgs))
    b = ctypes.c_void_p()
    assert lib.lbl_batch_create(labeler.handle, L.BLEQ, descs, len(imgs), 0, d_n.addr, 1, ctypes.byref(b)) == L.OK
    try:
        assert lib.lbl_batch_run(b, None) == L.OK
        hiprt.synchronize()
        counts = d_n.download((len(imgs),), np.uint32)
        for i, (img, (_, d_lab)) in enumerate(zip(imgs, bufs)):
            got = d_lab.download(img.shape, np.uint32)
            canon, m = O.bfs2d(img, 8)
            np.testing.assert_array_equal(O.canonical(got)[0], canon)
            assert counts[i] == m
    finally:
        lib.lbl_batch_destroy(b)


@pytest.mark.parametrize("flags", [0, 1, 2], ids=["graph", "autotune", "eager"])
def test_batch_canonical_labels(labeler, flags):
    """LBL_LABELS_CANONICAL on a batch: every image's labels are its own
    canonical numbering (1..n in raster order of first pixels), the oracle's,
    on every replay of the captured graph."""
    lib = labeler.lib
    imgs = _images(97, seed=23) + [np.zeros((5, 7), np.uint8), np.full((9, 4), 255, np.uint8)]
    bufs, descs = _upload(imgs, pitch_pad=flags)
    d_n = hiprt.DeviceBuffer(4 * len(imgs))
    b = ctypes.c_void_p()
    labeler.set_mode(L.LABELS_CANONICAL)
    try:
        st = lib.lbl_batch_create(labeler.handle, L.BLEQ, descs, len(imgs), 0, d_n.addr, flags, ctypes.byref(b))
        assert st == L.OK, labeler.lib.lbl_last_error(labeler.handle)
        print(f"[batch canonical flags={flags}]", flush=True)
        for rep in range(3):
            for _, d_lab in bufs:
                d_lab.zero()
            assert lib.lbl_batch_run(b, None) == L.OK
            hiprt.synchronize()
            counts = d_n.download((len(imgs),), np.uint32)
            for i, (img, (_, d_lab)) in enumerate(zip(imgs, bufs)):
                got = d_lab.download(img.shape, np.uint32)
                expect, m = O.bfs2d(img, 8)
                np.testing.assert_array_equal(got, expect, err_msg=f"image {i} {img.shape} replay {rep}")
                assert counts[i] == m
                if rep == 0 and i % 10 == 0:
                    single, n, _ = labeler.label2d(img, "bleq")
                    np.testing.assert_array_equal(single, got)
    finally:
        labeler.set_mode(L.LABELS_ROOT)
        if b.value:
            lib.lbl_batch_destroy(b)
