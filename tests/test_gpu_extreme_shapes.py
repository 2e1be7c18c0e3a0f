"""Images and volumes of extreme aspect ratio and large odd size, where runs,
bands and bars cross many tiles and the final kernel's spans wrap rows at every
alignment: one- and two-pixel-high rows a million pixels long, columns a
million pixels tall, odd sizes near 2^13 in each dimension; 3D needles and
slabs.  Every method's canonical labels and counts against the BFS labeler
(the reference's components, LabelingAlgs.cpp:336-529 for bleq)."""
import numpy as np
import pytest

import labeling_amd as L
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

SHAPES_2D = [(1, 1 << 20), (2, 200003), (3, 65537), (1 << 20, 1), (200003, 2), (65537, 3), (4097, 8193), (8191, 2049)]


@pytest.mark.parametrize("shape", SHAPES_2D, ids=lambda s: f"{s[0]}x{s[1]}")
def test_2d_extremes(labeler, shape):
    h, w = shape
    img = L.bernoulli(shape, 0.55, seed=h + 3 * w)
    print(f"[extreme 2D {h}x{w}]", flush=True)
    expect = {c: O.bfs2d(img, c) for c in (4, 8)}
    labeler.set_mode(L.LABELS_CANONICAL)
    try:
        for alg, conn, want in (("bleq", 0, 8), ("lbeq", 4, 4), ("lbeq", 8, 8), ("gr-block", 0, 8), ("he-run", 4, 4),
                                ("he-run", 8, 8), ("runeq", 0, 4), ("ocv", 4, 4)):
            lab, n, _ = labeler.label2d(img, alg, conn)
            ref, m = expect[want]
            assert n == m, (alg, conn)
            np.testing.assert_array_equal(lab, ref, err_msg=f"{alg} conn {conn}")
    finally:
        labeler.set_mode(L.LABELS_ROOT)
    lab, n, _ = labeler.label2d(img, "bleq")   # root values: 1 + minimum block index
    c, m = O.canonical(lab)
    np.testing.assert_array_equal(c, expect[8][0])
    assert n == m


SHAPES_3D = [(1, 1, 1 << 20), (1, 1 << 16, 3), (1 << 16, 2, 2), (3, 1025, 257), (257, 3, 1025), (129, 130, 131)]


@pytest.mark.parametrize("shape", SHAPES_3D, ids=lambda s: "x".join(map(str, s)))
def test_3d_extremes(labeler, shape):
    vol = L.bernoulli(shape, 0.3, seed=sum(shape))
    print(f"[extreme 3D {shape}]", flush=True)
    ref, m = O.bfs3d(vol, 26)
    for alg in ("bleq", "lbeq"):
        lab, n, _ = labeler.label3d(vol, alg)
        c, k = O.canonical(lab)
        np.testing.assert_array_equal(c, ref, err_msg=alg)
        assert n == m == k, alg
