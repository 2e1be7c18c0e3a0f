"""GPU canonical labels (LBL_LABELS_CANONICAL) of every path, through the C ABI.

In canonical mode the library itself numbers the components 1, 2, 3, ... in
raster order of their first pixel (voxel), x fastest, background 0.  Bar:
bit-exact.  The output must equal, without any host relabelling:
  * the scipy goldens (small fixtures; sha256 of the full-size workloads),
  * the oracle's breadth-first labelling (canonical by construction) on seeded
    random inputs of ragged shapes,
  * hand-built cases where a component's first pixel is not in its minimum
    block (the root that bleq's union-find keeps), which is what makes the
    device key pass necessary,
and the returned component count must equal the number of components.
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


@pytest.fixture
def canon(labeler):
    labeler.set_mode(L.LABELS_CANONICAL)
    yield labeler
    labeler.set_mode(L.LABELS_ROOT)


def _small(name, tags):
    z = np.load(os.path.join(GOLD, name))
    names = sorted({k.split("__")[0] for k in z.files})
    return [(n, z[f"{n}__img"], {t: z[f"{n}__{t}"] for t in tags}) for n in names]


@pytest.mark.parametrize("case", _small("small2d.npz", ["l4", "l8"]), ids=lambda c: c[0])
def test_small_goldens_2d(canon, case):
    name, img, gold = case
    for alg, conn, tag in [("bleq", 0, "l8"), ("lbeq", 8, "l8"), ("lbeq", 4, "l4")]:
        print(f"[canonical 2d golden {name} {alg} conn={conn}]", flush=True)
        lab, n, _ = canon.label2d(img, alg, conn=conn)
        np.testing.assert_array_equal(lab, gold[tag], err_msg=f"{alg} conn={conn}")
        assert n == int(gold[tag].max())


@pytest.mark.parametrize("alg", ["bleq", "lbeq"])
@pytest.mark.parametrize("case", _small("small3d.npz", ["l26"]), ids=lambda c: c[0])
def test_small_goldens_3d(canon, case, alg):
    name, vol, gold = case
    print(f"[canonical 3d golden {name} {alg}]", flush=True)
    lab, n, _ = canon.label3d(vol, alg)
    np.testing.assert_array_equal(lab, gold["l26"])
    assert n == int(gold["l26"].max())


SHAPES2 = [(1, 1), (2, 2), (1, 300), (300, 1), (3, 5), (31, 257), (33, 255), (65, 513), (257, 511), (517, 1029),
           (1024, 1024)]


@pytest.mark.parametrize("shape", SHAPES2, ids=lambda s: f"{s[0]}x{s[1]}")
@pytest.mark.parametrize("p", [0.2, 0.5, 0.8])
def test_random_2d(canon, shape, p):
    img = L.bernoulli(shape, p, seed=(shape[0] * 7919 + shape[1] * 31 + int(p * 10)) & 0xFFFF)
    for alg, conn in [("bleq", 0), ("lbeq", 4), ("lbeq", 8)]:
        print(f"[canonical 2d {shape} p={p} {alg} conn={conn}]", flush=True)
        lab, n, _ = canon.label2d(img, alg, conn=conn)
        expect, m = O.bfs2d(img, 8 if conn in (0, 8) else 4)
        np.testing.assert_array_equal(lab, expect, err_msg=f"{alg} conn={conn}")
        assert n == m


SHAPES3 = [(1, 1, 1), (2, 2, 2), (1, 1, 200), (200, 1, 1), (1, 200, 1), (3, 5, 7), (17, 17, 129), (33, 20, 70),
           (64, 64, 64), (40, 70, 260)]


@pytest.mark.parametrize("alg", ["bleq", "lbeq"])
@pytest.mark.parametrize("shape", SHAPES3, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("p", [0.1, 0.3, 0.6])
def test_random_3d(canon, shape, p, alg):
    vol = L.bernoulli(shape, p, seed=(shape[0] * 131 + shape[1] * 17 + shape[2] + int(p * 10)) & 0xFFFF)
    print(f"[canonical 3d {shape} p={p} {alg}]", flush=True)
    lab, n, _ = canon.label3d(vol, alg)
    expect, m = O.bfs3d(vol, 26)
    np.testing.assert_array_equal(lab, expect)
    assert n == m


def _first_pixel_not_in_min_block_2d():
    """Three components in block row 0 whose raster order differs from the
    order of their minimum blocks:
      C: (1,0) (2,1) (2,2..5) (1,6) (0,7): minimum block (0,0), first pixel (0,7)
      D: (0,20): block (10,0), first pixel (0,20)
      E: (1,12): block (6,0) -- before D's block -- but first pixel in row 1
    Canonical: C = 1, D = 2, E = 3."""
    img = np.zeros((4, 24), np.uint8)
    for y, x in [(1, 0), (2, 1), (2, 2), (2, 3), (2, 4), (2, 5), (1, 6), (0, 7)]:
        img[y, x] = 255
    img[0, 20] = 255
    img[1, 12] = 255
    expect = np.zeros(img.shape, np.uint32)
    for y, x in [(1, 0), (2, 1), (2, 2), (2, 3), (2, 4), (2, 5), (1, 6), (0, 7)]:
        expect[y, x] = 1
    expect[0, 20] = 2
    expect[1, 12] = 3
    return img, expect


def test_first_pixel_not_in_min_block_2d(canon):
    img, expect = _first_pixel_not_in_min_block_2d()
    for alg, conn in [("bleq", 0), ("lbeq", 8)]:
        lab, n, _ = canon.label2d(img, alg, conn=conn)
        np.testing.assert_array_equal(lab, expect, err_msg=alg)
        assert n == 3
    # the same pattern repeated across a wide image (many waves per block row)
    # and down it (every other block row), next to a solid band
    big = np.zeros((64, 24 * 100), np.uint8)
    for r in range(0, 64, 8):
        for c in range(100):
            big[r:r + 4, 24 * c:24 * c + 24] = img
    big[60:64, :] = 255
    lab, n, _ = canon.label2d(big, "bleq")
    expect_big, m = O.bfs2d(big, 8)
    np.testing.assert_array_equal(lab, expect_big)
    assert n == m


def test_first_voxel_not_in_min_block_3d(canon):
    """Component A has the minimum block (the origin block, through voxel
    (z=1, y=0, x=1)) but its first voxel in plane z=0 only at x=9; B, the single
    voxel (0, 0, 5), has a later minimum block (block column 2) yet comes
    first in raster order.  Canonical: B = 1, A = 2."""
    vol = np.zeros((4, 4, 16), np.uint8)
    for z, y, x in [(1, 0, 1), (1, 1, 2), (1, 2, 3), (1, 3, 4), (1, 3, 5), (1, 3, 6), (1, 2, 7), (1, 1, 8), (0, 0, 9)]:
        vol[z, y, x] = 255
    vol[0, 0, 5] = 255
    lab, n, _ = canon.label3d(vol, "bleq")
    expect, m = O.bfs3d(vol, 26)
    assert m == 2 and expect[0, 0, 5] == 1 and expect[0, 0, 9] == 2
    np.testing.assert_array_equal(lab, expect)
    assert n == 2
    for alg in ["bleq", "lbeq"]:
        rng = np.random.default_rng(5)
        vol = (rng.random((6, 10, 40)) < 0.08).astype(np.uint8) * 255
        lab, n, _ = canon.label3d(vol, alg)
        expect, m = O.bfs3d(vol, 26)
        np.testing.assert_array_equal(lab, expect, err_msg=alg)
        assert n == m


def test_empty_and_full(canon):
    for alg, conn in [("bleq", 0), ("lbeq", 4)]:
        lab, n, _ = canon.label2d(np.zeros((37, 45), np.uint8), alg, conn=conn)
        assert n == 0 and not lab.any()
        lab, n, _ = canon.label2d(np.full((37, 45), 255, np.uint8), alg, conn=conn)
        assert n == 1 and (lab == 1).all()
    for alg in ["bleq", "lbeq"]:
        lab, n, _ = canon.label3d(np.zeros((5, 6, 7), np.uint8), alg)
        assert n == 0 and not lab.any()
        lab, n, _ = canon.label3d(np.full((5, 6, 7), 255, np.uint8), alg)
        assert n == 1 and (lab == 1).all()


def test_modes_alternate(labeler):
    """Switching modes on one context reuses its workspace correctly."""
    img = L.bernoulli((300, 700), 0.45, seed=3)
    expect, m = O.bfs2d(img, 8)
    for mode in [L.LABELS_ROOT, L.LABELS_CANONICAL, L.LABELS_ROOT, L.LABELS_CANONICAL]:
        labeler.set_mode(mode)
        try:
            lab, n, _ = labeler.label2d(img, "bleq")
        finally:
            labeler.set_mode(L.LABELS_ROOT)
        assert n == m
        if mode == L.LABELS_CANONICAL:
            np.testing.assert_array_equal(lab, expect)
        else:
            np.testing.assert_array_equal(O.canonical(lab)[0], expect)


def _large():
    path = os.path.join(GOLD, "large.json")
    return json.load(open(path)) if os.path.exists(path) else {}


@pytest.mark.slow
@pytest.mark.parametrize("name", [k for k in sorted(_large()) if k[:3] in ("c1_", "c2_")])
def test_full_size_2d(canon, name):
    entry = _large()[name]
    spec = entry["spec"]
    if spec["kind"] == "bernoulli":
        img = L.bernoulli(tuple(spec["shape"]), spec["p"], spec["seed"])
    else:
        img = L.blobs2d(*spec["shape"], spec["sigma"], spec["q"], spec["seed"])
    assert hashlib.sha256(img.tobytes()).hexdigest() == entry["input_sha256"]
    runs = [("bleq", 0, "conn8")] + ([("lbeq", 8, "conn8"), ("lbeq", 4, "conn4")] if "conn4" in entry else [])
    for alg, conn, tag in runs:
        print(f"[canonical 2d full {name} {alg} conn={conn}]", flush=True)
        lab, n, _ = canon.label2d(img, alg, conn=conn)
        assert hashlib.sha256(lab.tobytes()).hexdigest() == entry[tag]["labels_sha256"]
        assert n == entry[tag]["components"]


@pytest.mark.slow
@pytest.mark.parametrize("alg", ["bleq", "lbeq"])
@pytest.mark.parametrize("name", [k for k in sorted(_large()) if k.startswith("c3_")])
def test_full_size_3d(canon, name, alg):
    entry = _large()[name]
    spec = entry["spec"]
    print(f"[canonical 3d full {name} {alg}]", flush=True)
    if spec["kind"] == "bernoulli":
        vol = L.bernoulli(tuple(spec["shape"]), spec["p"], spec["seed"])
    else:
        vol = L.blobs3d(*spec["shape"], spec["sigma"], spec["q"], spec["seed"])
    assert hashlib.sha256(vol.tobytes()).hexdigest() == entry["input_sha256"]
    lab, n, _ = canon.label3d(vol, alg)
    del vol
    assert hashlib.sha256(lab.tobytes()).hexdigest() == entry["conn26"]["labels_sha256"]
    assert n == entry["conn26"]["components"]
