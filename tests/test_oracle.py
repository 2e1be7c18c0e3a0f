"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Checks, chosen so that a dropped term, a wrong offset or a transposed axis in the
restatement fails at least one of them:
  * exhaustive: every 4x4, 4x5 and 5x4 image (bleq 8-conn, lbeq 4- and 8-conn)
    and every 2x2x2, 2x2x3 and 2x3x3 volume (bleq / lbeq 26-conn) agrees with
    the independent BFS labeler after canonical relabelling;
  * golden fixtures written by scipy.ndimage.label (tests/golden/);
  * known-answer images derived from the reference source;
  * invariants: block connectivity bits are symmetric, raw labels are
    1 + the minimum padded block / pixel index of each component.
"""
import json
import os

import numpy as np
import pytest

import labeling_amd as L
from oracle import pyoracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _small2d():
    z = np.load(os.path.join(GOLD, "small2d.npz"))
    names = sorted({k.split("__")[0] for k in z.files})
    return [(n, z[f"{n}__img"], z[f"{n}__l8"], z[f"{n}__l4"]) for n in names]


def _small3d():
    z = np.load(os.path.join(GOLD, "small3d.npz"))
    names = sorted({k.split("__")[0] for k in z.files})
    return [(n, z[f"{n}__img"], z[f"{n}__l26"]) for n in names]


# ---- exhaustive ------------------------------------------------------------------

@pytest.mark.parametrize("h,w", [(4, 4), (4, 5), (5, 4), (1, 7), (7, 1), (3, 5)])
def test_exhaustive_bleq2d_padded(h, w):
    bad, first = O.exhaustive2d(h, w, 0, 8)
    assert bad == 0, f"first failing image code {first}"


@pytest.mark.parametrize("h,w,conn", [(4, 4, 4), (4, 5, 4), (5, 4, 4), (4, 4, 8), (4, 5, 8), (5, 4, 8)])
def test_exhaustive_lbeq2d(h, w, conn):
    bad, first = O.exhaustive2d(h, w, 1, conn)
    assert bad == 0, f"first failing image code {first}"


@pytest.mark.parametrize("shape,alg", [((2, 2, 2), 0), ((2, 2, 3), 0), ((2, 3, 3), 0), ((3, 3, 2), 0),
                                       ((3, 2, 3), 0), ((2, 2, 2), 1), ((2, 3, 3), 1)])
def test_exhaustive_3d(shape, alg):
    bad, first = O.exhaustive3d(*shape, alg)
    assert bad == 0, f"first failing volume code {first}"


def test_unpadded_reference_cpu_is_not_exact():
    """The reference CPU bleq on an unpadded image (LabelingAlgs.cpp:353-376) wraps
    at the right edge; the exhaustive sweep must find it, first at KAT-1's image."""
    bad, first = O.exhaustive2d(4, 4, 2, 8)
    assert bad > 0
    assert first == (1 << 3) | (1 << 4)   # pixels (x=3,y=0) and (x=0,y=1)


# ---- golden fixtures ------------------------------------------------------------------

@pytest.mark.parametrize("case", _small2d(), ids=lambda c: c[0])
def test_golden_small2d(case):
    name, img, l8, l4 = case
    a, n = O.canonical(O.bleq2d(img))
    assert n == int(l8.max())
    np.testing.assert_array_equal(a, l8)
    for conn, ref in ((8, l8), (4, l4)):
        b, _ = O.canonical(O.lbeq2d(img, conn=conn))
        np.testing.assert_array_equal(b, ref)
        c, m = O.bfs2d(img, conn)
        np.testing.assert_array_equal(c, ref)
        assert m == int(ref.max())


@pytest.mark.parametrize("case", _small3d(), ids=lambda c: c[0])
def test_golden_small3d(case):
    name, vol, l26 = case
    a, n = O.canonical(O.bleq3d(vol))
    np.testing.assert_array_equal(a, l26)
    b, _ = O.canonical(O.lbeq3d(vol))
    np.testing.assert_array_equal(b, l26)
    c, m = O.bfs3d(vol, 26)
    np.testing.assert_array_equal(c, l26)
    assert n == m == int(l26.max())


def _large():
    path = os.path.join(GOLD, "large.json")
    return json.load(open(path)) if os.path.exists(path) else {}


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_golden_c1_1024(seed):
    """1024x1024 p=0.5 (the CPU-reference workload): input bytes and labels vs scipy."""
    import hashlib

    entry = _large()[f"c1_bern1024_p5_s{seed}"]
    img = L.bernoulli((1024, 1024), 0.5, seed)
    assert hashlib.sha256(img.tobytes()).hexdigest() == entry["input_sha256"]
    a, n = O.canonical(O.bleq2d(img))
    assert hashlib.sha256(a.tobytes()).hexdigest() == entry["conn8"]["labels_sha256"]
    assert n == entry["conn8"]["components"]
    b, m = O.canonical(O.lbeq2d(img, conn=4))
    assert hashlib.sha256(b.tobytes()).hexdigest() == entry["conn4"]["labels_sha256"]
    assert m == entry["conn4"]["components"]


# ---- known answers (derived from the reference source) --------------------------------

def test_kat1_right_edge_wrap():
    img = np.zeros((4, 4), np.uint8)
    img[0, 3] = img[1, 0] = 255
    exact, n = O.canonical(O.bleq2d(img, padded=True))
    assert n == 2
    raw, info = O.bleq2d(img, padded=False, with_info=True)
    _, m = O.canonical(raw)
    assert m == 1 and raw[0, 3] == raw[1, 0] == 1     # the reference CPU merges them
    assert not info["oob"]


def test_kat2_corner_adjacency():
    img = np.zeros((4, 4), np.uint8)
    img[1, 1] = img[2, 2] = 255
    assert O.canonical(O.bleq2d(img))[1] == 1
    assert O.canonical(O.lbeq2d(img, conn=0))[1] == 2     # lbeq default is 4-connectivity
    assert O.canonical(O.lbeq2d(img, conn=8))[1] == 1
    vol = np.zeros((4, 4, 4), np.uint8)
    vol[1, 1, 1] = vol[2, 2, 2] = 255
    assert O.canonical(O.bleq3d(vol))[1] == 1
    assert O.canonical(O.lbeq3d(vol))[1] == 1


def test_kat3_pixel_origin():
    img = np.zeros((3, 3), np.uint8)
    img[0, 0] = img[0, 1] = img[1, 0] = 255
    assert O.canonical(O.lbeq2d(img, conn=4))[1] == 1
    q = O.lbeq2d(img, conn=4, quirks=True)          # reference CPU: 0-based, pos 0 excluded
    assert q[0, 0] == 0 and q[0, 1] == 1 and q[1, 0] == 3
    assert O.canonical(q)[1] == 2


def test_kat4_serpentine_iterations():
    img = np.zeros((33, 41), np.uint8)
    for r in range(0, 33, 2):
        img[r, :] = 255
        if r + 1 < 33:
            img[r + 1, 40 if (r // 2) % 2 == 0 else 0] = 255
    lab, info = O.bleq2d(img, with_info=True)
    assert O.canonical(lab)[1] == 1
    assert 2 <= info["iterations"] < 100
    lab, info = O.lbeq2d(img, conn=4, with_info=True)
    assert O.canonical(lab)[1] == 1


@pytest.mark.parametrize("shape", [(1, 1), (1, 9), (9, 1), (31, 31), (33, 63), (2, 2)])
def test_kat5_degenerate(shape):
    for img in (np.zeros(shape, np.uint8), np.full(shape, 255, np.uint8)):
        a, n = O.canonical(O.bleq2d(img))
        assert n == int(img.any())
        b, m = O.canonical(O.lbeq2d(img, conn=4))
        assert m == int(img.any())
    yy, xx = np.mgrid[0:shape[0], 0:shape[1]]
    chk = np.where((yy + xx) % 2 == 0, 255, 0).astype(np.uint8)
    single_line = min(shape) == 1
    assert O.canonical(O.bleq2d(chk))[1] == (int(np.count_nonzero(chk)) if single_line else 1)
    assert O.canonical(O.lbeq2d(chk, conn=4))[1] == int(np.count_nonzero(chk))


# ---- invariants ---------------------------------------------------------------------

def test_conn_bits_symmetric_2d():
    img = L.bernoulli((61, 77), 0.5, 3)
    c = O.bleq2d_conn(img).astype(np.int64)
    dxy = [(-1, -1), (0, -1), (1, -1), (-1, 0), (1, 0), (-1, 1), (0, 1), (1, 1)]
    bh, bw = c.shape
    for k, (dx, dy) in enumerate(dxy):
        mirror = dxy.index((-dx, -dy))
        for by in range(1, bh - 1):
            for bx in range(1, bw - 1):
                assert ((c[by, bx] >> k) & 1) == ((c[by + dy, bx + dx] >> mirror) & 1)


def test_conn_bits_symmetric_3d():
    vol = L.bernoulli((9, 11, 10), 0.3, 4)
    c = O.bleq3d_conn(vol).astype(np.int64)
    s = c.shape
    for bit in range(27):
        if bit == 13:
            continue
        d0, d1, d2 = bit % 3 - 1, (bit // 3) % 3 - 1, bit // 9 - 1
        mirror = (1 - d0) + 3 * (1 - d1) + 9 * (1 - d2)
        for x in range(1, s[0] - 1):
            for y in range(1, s[1] - 1):
                for z in range(1, s[2] - 1):
                    assert ((c[x, y, z] >> bit) & 1) == ((c[x + d0, y + d1, z + d2] >> mirror) & 1)


def test_raw_labels_are_min_block_index():
    """GPU raw value of the reference = 1 + min padded-block index (LabelingAlgs.cl:202,317)."""
    img = L.bernoulli((45, 70), 0.4, 5)
    raw = O.bleq2d(img)
    canon, n = O.bfs2d(img, 8)
    pbw = O.load().orc_pad2(70) // 2
    for k in range(1, n + 1):
        ys, xs = np.nonzero(canon == k)
        expect = 1 + np.min((ys // 2) * pbw + xs // 2)
        assert set(np.unique(raw[canon == k])) == {expect}
    raw4 = O.lbeq2d(img, conn=4)
    canon4, n4 = O.bfs2d(img, 4)
    pw = O.load().orc_pad2(70)
    for k in range(1, n4 + 1):
        ys, xs = np.nonzero(canon4 == k)
        assert set(np.unique(raw4[canon4 == k])) == {1 + np.min(ys * pw + xs)}


def test_canonicalize_matches_numpy():
    rng = np.random.default_rng(0)
    lab = rng.integers(0, 50, size=5000).astype(np.uint32)
    a, n = O.canonical(lab)
    vals, first = np.unique(lab[lab != 0], return_index=True)
    assert n == len(vals)
    order = vals[np.argsort(first)]
    expect = np.zeros_like(lab)
    for i, v in enumerate(order):
        expect[lab == v] = i + 1
    np.testing.assert_array_equal(a, expect)


@pytest.mark.parametrize("conn", [4, 8])
@pytest.mark.parametrize("shape,p,origin", [((37, 53), 0.55, 255), ((64, 64), 0.45, 0), ((5, 33), 0.7, 255),
                                            ((1, 40), 0.8, 255), ((40, 1), 0.8, 255)])
def test_lbeq2d_gpu_raw_values_closed_form(conn, shape, p, origin):
    """orc_lbeq2d quirks=2 (the reference GPU kernels' raw values) equals a
    closed form built from the independent BFS: drop pixel (0,0), label each
    remaining component by its first pixel's index in the padded image
    (0-based); pixel (0,0) keeps 0."""
    img = L.bernoulli(shape, p, seed=sum(shape) + conn)
    img[0, 0] = origin
    got = O.lbeq2d(img, conn=conn, quirks=2)
    dropped = img.copy()
    dropped[0, 0] = 0
    lab, n = O.bfs2d(dropped, conn)
    wp = O.load().orc_pad2(shape[1])
    ys, xs = np.nonzero(lab)
    first = np.full(n + 1, np.iinfo(np.int64).max)
    np.minimum.at(first, lab[ys, xs], ys.astype(np.int64) * wp + xs)
    expect = np.zeros(shape, np.int64)
    expect[ys, xs] = first[lab[ys, xs]]
    np.testing.assert_array_equal(got, expect.astype(np.uint32))
    assert got[0, 0] == 0
