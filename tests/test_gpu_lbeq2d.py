"""GPU parity of 2D label equivalence (lbeq) against the oracle, through the C ABI.

Bit-exact: canonical labels and counts equal the oracle / scipy goldens for
4- and 8-connectivity, and the raw labels equal 1 + the minimum pixel index of
each component (the oracle's restatement of the reference with 1-based
labels: 1 + min padded index, re-expressed in the unpadded raster).
"""
import hashlib
import json
import os

import numpy as np
import pytest

import labeling_amd as L
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _small2d():
    z = np.load(os.path.join(GOLD, "small2d.npz"))
    names = sorted({k.split("__")[0] for k in z.files})
    return [(n, z[f"{n}__img"], z[f"{n}__l8"], z[f"{n}__l4"]) for n in names]


def _min_pixel_labels(img, canon, n):
    ys, xs = np.nonzero(img)
    w = img.shape[1]
    comp = canon[ys, xs]
    mins = np.full(n + 1, np.iinfo(np.int64).max)
    np.minimum.at(mins, comp, ys.astype(np.int64) * w + xs)
    out = np.zeros(img.shape, np.int64)
    out[ys, xs] = mins[comp] + 1
    return out


@pytest.mark.parametrize("case", _small2d(), ids=lambda c: c[0])
def test_small_goldens(labeler, case):
    name, img, l8, l4 = case
    print(f"[lbeq2d golden {name} {img.shape}]", flush=True)
    for conn, ref in ((4, l4), (8, l8), (0, l4)):
        lab, n, _ = labeler.label2d(img, "lbeq", conn=conn)
        c, m = O.canonical(lab)
        np.testing.assert_array_equal(c, ref)
        assert n == m == int(ref.max())


SHAPES = [(1, 1), (1, 300), (300, 1), (3, 5), (31, 127), (32, 128), (33, 129), (64, 256), (65, 257),
          (257, 511), (100, 1000), (1000, 100), (517, 1029), (1024, 1024)]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"{s[0]}x{s[1]}")
@pytest.mark.parametrize("conn", [4, 8])
def test_vs_oracle_and_root_values(labeler, shape, conn):
    img = L.bernoulli(shape, 0.55, seed=31 + conn)
    print(f"[lbeq2d {shape} conn={conn}]", flush=True)
    lab, n, _ = labeler.label2d(img, "lbeq", conn=conn)
    expect, m = O.canonical(O.lbeq2d(img, conn=conn))
    c, k = O.canonical(lab)
    np.testing.assert_array_equal(c, expect)
    assert n == m == k
    canon, _ = O.bfs2d(img, conn)
    np.testing.assert_array_equal(lab.astype(np.int64), _min_pixel_labels(img, canon, m))


@pytest.mark.parametrize("p", [0.1, 0.5, 0.9])
def test_bernoulli_mid_size(labeler, p):
    img = L.bernoulli((2000, 3001), p, seed=7)
    for conn in (4, 8):
        lab, n, _ = labeler.label2d(img, "lbeq", conn=conn)
        canon, m = O.bfs2d(img, conn)
        np.testing.assert_array_equal(O.canonical(lab)[0], canon)
        assert n == m


def test_serpentine_and_iteration_cap(labeler):
    ser = np.zeros((301, 999), np.uint8)
    for r in range(0, 301, 2):
        ser[r, :] = 255
        if r + 1 < 301:
            ser[r + 1, 998 if (r // 2) % 2 == 0 else 0] = 255
    lab, n, _ = labeler.label2d(ser, "lbeq", conn=4)
    assert n == 1 and O.canonical(lab)[1] == 1
    labeler.set_iteration_cap(1)
    try:
        with pytest.raises(L.LabelingError) as e:
            labeler.label2d(ser, "lbeq", conn=4)
        assert e.value.status == L.EITER
    finally:
        labeler.set_iteration_cap(0)
    lab, n, _ = labeler.label2d(ser, "lbeq", conn=4)   # the context recovers
    assert n == 1


def test_reference_mode_rejected_for_lbeq(labeler):
    labeler.set_mode(L.LABELS_REFERENCE)
    try:
        with pytest.raises(L.LabelingError) as e:
            labeler.label2d(np.zeros((4, 4), np.uint8), "lbeq")
        assert e.value.status == L.EUNSUPPORTED
    finally:
        labeler.set_mode(L.LABELS_ROOT)


@pytest.mark.slow
def test_full_size_c1_conn4(labeler):
    entry = json.load(open(os.path.join(GOLD, "large.json")))["c1_bern1024_p5_s0"]
    img = L.bernoulli((1024, 1024), 0.5, 0)
    lab, n, _ = labeler.label2d(img, "lbeq", conn=4)
    c, m = O.canonical(lab)
    assert hashlib.sha256(c.tobytes()).hexdigest() == entry["conn4"]["labels_sha256"]
    assert n == m == entry["conn4"]["components"]


@pytest.mark.slow
@pytest.mark.parametrize("name", ["c2_bern8192_p5", "c2_blob8192"])
def test_full_size_conn8(labeler, name):
    entry = json.load(open(os.path.join(GOLD, "large.json")))[name]
    spec = entry["spec"]
    if spec["kind"] == "bernoulli":
        img = L.bernoulli(tuple(spec["shape"]), spec["p"], spec["seed"])
    else:
        img = L.blobs2d(*spec["shape"], spec["sigma"], spec["q"], spec["seed"])
    lab, n, _ = labeler.label2d(img, "lbeq", conn=8)
    c, m = O.canonical(lab)
    assert hashlib.sha256(c.tobytes()).hexdigest() == entry["conn8"]["labels_sha256"]
    assert n == m == entry["conn8"]["components"]


@pytest.mark.parametrize("shape", [(257, 511), (1000, 100), (517, 1029)], ids=lambda s: f"{s[0]}x{s[1]}")
def test_block_path_equals_sweep_path(labeler, shape):
    """4-connectivity runs on 1x2 blocks; with an iteration cap set, it runs the
    tile sweeps instead.  Both give the same raw labels."""
    img = L.bernoulli(shape, 0.5, seed=99)
    print(f"[lbeq2d block vs sweeps {shape}]", flush=True)
    fast, n1, _ = labeler.label2d(img, "lbeq", conn=4)
    labeler.set_iteration_cap(1 << 30)
    try:
        swept, n2, _ = labeler.label2d(img, "lbeq", conn=4)
    finally:
        labeler.set_iteration_cap(0)
    np.testing.assert_array_equal(fast, swept)
    assert n1 == n2
