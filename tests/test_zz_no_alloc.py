"""lbl_label2d_dev_noalloc (the entry point that allocates nothing; its
workspace is a fixed per-device scratch array) against lbl_label2d_dev on the
same inputs: identical labels and counts, on calls in a row whose sizes shrink
(the scratch keeps the previous call's contents), and a clear error above the
scratch's limit."""
import ctypes

import numpy as np
import pytest

import labeling_amd as L
from tests import hiprt

pytestmark = pytest.mark.gpu


def _both(labeler, img, alg, conn, with_count=True):
    h, w = img.shape
    d_px = hiprt.to_device(img)
    d_a, d_b = hiprt.DeviceBuffer(4 * h * w), hiprt.DeviceBuffer(4 * h * w)
    d_na, d_nb = hiprt.DeviceBuffer(4), hiprt.DeviceBuffer(4)
    d_a.upload(np.full((h, w), 0xABCDEF, np.uint32))
    d_b.upload(np.full((h, w), 0x123456, np.uint32))
    labeler.label2d_dev(d_px.addr, h, w, w, d_a.addr, alg, conn, d_na.addr)
    st = labeler.lib.lbl_label2d_dev_noalloc(labeler.handle, L.ALGS[alg], d_px.addr, h, w, w, conn, d_b.addr,
                                             d_nb.addr if with_count else None, None)
    assert st == L.OK, labeler.lib.lbl_last_error(labeler.handle)
    hiprt.synchronize()
    a, b = d_a.download((h, w), np.uint32), d_b.download((h, w), np.uint32)
    na, nb = int(d_na.download((1,), np.uint32)[0]), int(d_nb.download((1,), np.uint32)[0])
    return a, b, na, nb


@pytest.mark.parametrize("alg,conn", [("bleq", 0), ("lbeq", 4), ("lbeq", 8)])
def test_noalloc_equals_dev_two_calls_shrinking(labeler, alg, conn):
    big = (8192, 8192) if alg == "bleq" else (1500, 2047)
    for k, (shape, p) in enumerate([(big, 0.5), ((777, 1031), 0.3)]):
        img = L.bernoulli(shape, p, seed=40 + k)
        print(f"[noalloc {alg} conn={conn} call {k} {shape}]", flush=True)
        a, b, na, nb = _both(labeler, img, alg, conn)
        np.testing.assert_array_equal(b, a)
        assert na == nb


def test_noalloc_count_after_uncounted_call(labeler):
    """A call without a count pointer leaves the scratch's count slots as they
    were; the next counted call still counts from zero."""
    img = L.bernoulli((300, 500), 0.5, seed=3)
    _both(labeler, img, "bleq", 0, with_count=False)
    a, b, na, nb = _both(labeler, img, "bleq", 0)
    np.testing.assert_array_equal(b, a)
    assert na == nb


def test_noalloc_rejects_beyond_scratch(labeler):
    h, w = 8690, 8690   # 4345^2 2x2 blocks: block words just beyond the 72 MiB scratch (18,873,344 blocks)
    d_px = hiprt.DeviceBuffer(h * w)
    d_lab = hiprt.DeviceBuffer(4 * h * w)
    st = labeler.lib.lbl_label2d_dev_noalloc(labeler.handle, L.BLEQ, d_px.addr, h, w, w, 0, d_lab.addr, None, None)
    assert st == L.ENOMEM
    assert b"workspace" in labeler.lib.lbl_last_error(labeler.handle)
    h, w = 8688, 8688   # 4344^2 blocks: fits
    img = L.bernoulli((h, w), 0.5, seed=9)
    a, b, na, nb = _both(labeler, img, "bleq", 0)
    np.testing.assert_array_equal(b, a)
    assert na == nb
