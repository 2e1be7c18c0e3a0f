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


def _unpadded_values(raw, w):
    """Oracle labels 1 + min padded pixel index (padded width orc_pad2(w)) as
    1 + the same pixel's index in the unpadded raster."""
    wp = L.ref_pad2(w)
    v = raw.astype(np.int64)
    out = np.where(v > 0, ((v - 1) // wp) * w + (v - 1) % wp + 1, 0)
    return out.astype(np.uint32)


SWEEP_SHAPES = [(1, 1), (3, 5), (31, 127), (33, 129), (257, 511), (1000, 100), (517, 1029)]


@pytest.mark.parametrize("shape", SWEEP_SHAPES, ids=lambda s: f"{s[0]}x{s[1]}")
@pytest.mark.parametrize("conn", [4, 8])
@pytest.mark.parametrize("p", [0.3, 0.6])
def test_sweep_path_vs_oracle(labeler, shape, conn, p):
    """The Kalentev sweep kernels (lbeq2d_tile / merge / flatten / final, run
    when an iteration cap is set) against the oracle's restatement of
    TLabelDistribution (LabelingAlgs.cpp:211-330) / DistrScanKernel
    (LabelingAlgs.cl:86-130) directly: raw values bit-exact, and equal to the
    block-forest path's."""
    img = L.bernoulli(shape, p, seed=hash((shape, conn, p)) & 0xFFFF)
    print(f"[lbeq2d sweeps vs oracle {shape} conn={conn} p={p}]", flush=True)
    expect = _unpadded_values(O.lbeq2d(img, conn=conn), shape[1])
    fast, n1, _ = labeler.label2d(img, "lbeq", conn=conn)
    labeler.set_iteration_cap(1 << 30)
    try:
        swept, n2, _ = labeler.label2d(img, "lbeq", conn=conn)
    finally:
        labeler.set_iteration_cap(0)
    np.testing.assert_array_equal(swept, expect)
    np.testing.assert_array_equal(fast, expect)
    assert n1 == n2 == O.canonical(expect)[1]


def test_sweep_path_serpentine_and_small_goldens(labeler):
    """The worst case for label propagation, and the small scipy goldens, on
    the sweep path."""
    ser = np.zeros((301, 999), np.uint8)
    for r in range(0, 301, 2):
        ser[r, :] = 255
        if r + 1 < 301:
            ser[r + 1, 998 if (r // 2) % 2 == 0 else 0] = 255
    labeler.set_iteration_cap(1 << 30)
    try:
        for conn in (4, 8):
            lab, n, _ = labeler.label2d(ser, "lbeq", conn=conn)
            np.testing.assert_array_equal(lab, _unpadded_values(O.lbeq2d(ser, conn=conn), 999))
            assert n == 1
        for name, img, l8, l4 in _small2d():
            for conn, ref in ((4, l4), (8, l8)):
                lab, n, _ = labeler.label2d(img, "lbeq", conn=conn)
                np.testing.assert_array_equal(O.canonical(lab)[0], ref, err_msg=f"{name} conn={conn}")
                assert n == int(ref.max())
    finally:
        labeler.set_iteration_cap(0)


@pytest.mark.slow
@pytest.mark.parametrize("name", ["c2_bern8192_p1", "c2_bern8192_p5", "c2_bern8192_p9", "c2_blob8192"])
def test_full_size_conn4_8192(labeler, name):
    """8192^2 inputs with 4-connectivity (the lbeq default) against scipy goldens."""
    entry = json.load(open(os.path.join(GOLD, "large.json")))[name]
    spec = entry["spec"]
    if spec["kind"] == "bernoulli":
        img = L.bernoulli(tuple(spec["shape"]), spec["p"], spec["seed"])
    else:
        img = L.blobs2d(*spec["shape"], spec["sigma"], spec["q"], spec["seed"])
    print(f"[lbeq2d conn4 {name}]", flush=True)
    lab, n, _ = labeler.label2d(img, "lbeq", conn=4)
    c, m = O.canonical(lab)
    assert hashlib.sha256(c.tobytes()).hexdigest() == entry["conn4"]["labels_sha256"]
    assert n == m == entry["conn4"]["components"]
