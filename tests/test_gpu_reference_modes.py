"""LBL_LABELS_REFERENCE for every path, and LBL_LABELS_REFERENCE_UNCROPPED for
3D: the reference OpenCL implementation's raw output values, bit-exact against
the oracle's restatement of its kernels.

  2D lbeq  min padded pixel index, 0-based, pixel (0,0) never labelled
           (DistrInitKernel / GetLabel / DistrScanKernel, LabelingAlgs.cl:31-110;
           oracle quirks=2)
  3D bleq  1 + min block index of the padded block grid (BLEQ3D_Init,
           LabelingAlgs.cl:828; oracle orc_bleq3d)
  3D lbeq  1 + min padded voxel index (LBEQ3D_InitKernel, LabelingAlgs.cl:641;
           oracle orc_lbeq3d)
  3D uncropped: the padded volume IOCLLabeling3D::Label returns
           (LabelingTools.cpp:160-221), data at +2, zeros elsewhere.
(2D bleq's raw values are tested in test_gpu_bleq2d.py.)
"""
import numpy as np
import pytest

import labeling_amd as L
from oracle import pyoracle as O
from tests import hiprt

pytestmark = pytest.mark.gpu


@pytest.fixture()
def mode(labeler):
    def set_mode(m):
        labeler.set_mode(m)
        return labeler
    yield set_mode
    labeler.set_mode(L.LABELS_ROOT)


SHAPES2D = [(1, 1), (1, 77), (77, 1), (2, 2), (3, 5), (31, 127), (33, 129), (64, 256), (257, 511), (517, 1029),
            (1024, 1024)]


@pytest.mark.parametrize("shape", SHAPES2D, ids=lambda s: f"{s[0]}x{s[1]}")
@pytest.mark.parametrize("conn", [4, 8])
@pytest.mark.parametrize("p", [0.3, 0.6])
def test_lbeq2d_reference_values(mode, shape, conn, p):
    img = L.bernoulli(shape, p, seed=hash((shape, conn, p)) & 0xFFFF)
    img[0, 0] = 255 if p > 0.5 else 0   # both cases of the unlabelled origin
    print(f"[lbeq2d ref {shape} conn={conn} p={p}]", flush=True)
    lab, n, _ = mode(L.LABELS_REFERENCE).label2d(img, "lbeq", conn=conn)
    expect = O.lbeq2d(img, conn=conn, quirks=2)
    np.testing.assert_array_equal(lab, expect)
    assert n == O.canonical(expect)[1]


def test_lbeq2d_reference_origin_splits(mode):
    """Pixel (0,0) joins (1,0) and (0,1) under 4-connectivity; the reference
    never labels it, so its two neighbours become two components."""
    img = np.zeros((3, 3), np.uint8)
    img[0, 0] = img[0, 1] = img[1, 0] = 255
    lab, n, _ = mode(L.LABELS_REFERENCE).label2d(img, "lbeq", conn=4)
    wp = L.ref_pad2(3)
    np.testing.assert_array_equal(lab, np.array([[0, 1, 0], [wp, 0, 0], [0, 0, 0]], np.uint32))
    assert n == 2
    lab8, n8, _ = mode(L.LABELS_REFERENCE).label2d(img, "lbeq", conn=8)   # (0,1)-(1,0) touch diagonally
    np.testing.assert_array_equal(lab8, np.array([[0, 1, 0], [1, 0, 0], [0, 0, 0]], np.uint32))
    assert n8 == 1


SHAPES3D = [(1, 1, 1), (2, 3, 5), (9, 13, 11), (16, 16, 16), (33, 17, 65), (64, 64, 64), (30, 70, 129)]


@pytest.mark.parametrize("shape", SHAPES3D, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("alg", ["bleq", "lbeq"])
def test_3d_reference_values(mode, shape, alg):
    vol = L.bernoulli(shape, 0.3, seed=hash((shape, alg)) & 0xFFFF)
    print(f"[3d ref {alg} {shape}]", flush=True)
    expect = O.bleq3d(vol) if alg == "bleq" else O.lbeq3d(vol)
    lab, n, _ = mode(L.LABELS_REFERENCE).label3d(vol, alg)
    np.testing.assert_array_equal(lab, expect)
    assert n == O.canonical(expect)[1]
    pad, n2, _ = mode(L.LABELS_REFERENCE_UNCROPPED).label3d(vol, alg, padded=True)
    d, h, w = shape
    assert pad.shape == (L.ref_pad3(d), L.ref_pad3(h), L.ref_pad3(w))
    full = np.zeros(pad.shape, np.uint32)
    full[2:d + 2, 2:h + 2, 2:w + 2] = expect
    np.testing.assert_array_equal(pad, full)
    assert n2 == n


@pytest.mark.parametrize("alg", ["bleq", "lbeq"])
def test_3d_uncropped_device_entry_point(mode, alg):
    """lbl_label3d_dev writes the padded volume into a caller buffer of that
    size, every element (the padding too: the buffer starts as garbage)."""
    vol = L.blobs3d(40, 50, 70, 3.0, 0.4, seed=5)
    d, h, w = vol.shape
    pshape = (L.ref_pad3(d), L.ref_pad3(h), L.ref_pad3(w))
    d_in = hiprt.to_device(vol)
    d_lab = hiprt.DeviceBuffer(4 * int(np.prod(pshape)))
    d_lab.upload(np.full(pshape, 0xDEADBEEF, np.uint32))
    lab = mode(L.LABELS_REFERENCE_UNCROPPED)
    lab.label3d_dev(d_in.addr, d, h, w, d_lab.addr, alg)
    hiprt.synchronize()
    got = d_lab.download(pshape, np.uint32)
    full = np.zeros(pshape, np.uint32)
    full[2:d + 2, 2:h + 2, 2:w + 2] = O.bleq3d(vol) if alg == "bleq" else O.lbeq3d(vol)
    np.testing.assert_array_equal(got, full)


def test_mode_errors(mode):
    lab = mode(L.LABELS_REFERENCE_UNCROPPED)
    with pytest.raises(L.LabelingError) as e:
        lab.label2d(np.zeros((4, 4), np.uint8), "bleq")
    assert e.value.status == L.EUNSUPPORTED
    lab = mode(L.LABELS_REFERENCE)
    lab.set_iteration_cap(10)
    try:
        with pytest.raises(L.LabelingError) as e:
            lab.label2d(np.zeros((4, 4), np.uint8), "lbeq")
        assert e.value.status == L.EUNSUPPORTED
        lab.label2d(np.zeros((4, 4), np.uint8), "bleq")   # bleq has no sweep path: unaffected
    finally:
        lab.set_iteration_cap(0)
