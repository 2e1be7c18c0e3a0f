"""The oracle's restatements of the reference registry's other methods
(oracle/oracle_registry.c), pinned on the CPU:

* components: every restatement equals the independent BFS labeler
  (4-connectivity for run equivalence and He's 4-connected mode, 8 for He's
  8-connected mode and Grana's block labeling);
* values: equal to the closed forms of tests/registry_values.py (first-run
  numbering, provisional-label ranks, block order), exhaustively on tiny images
  and on random ones;
* the reference's quirks (He's labeler skips the last row and the last
  column's runs, LabelingAlgs.cpp:150-171): the quirk-free restatement equals
  the as-written one wherever those rows and columns are empty, and a known
  answer shows the difference.
"""
import itertools

import numpy as np
import pytest

from oracle import pyoracle as O
from tests import registry_values as V


def _all_images(h, w):
    for code in range(1 << (h * w)):
        yield np.array([(code >> i) & 1 for i in range(h * w)], np.uint8).reshape(h, w)


def _random_images(n, seed, hi=14):
    rng = np.random.default_rng(seed)
    for _ in range(n):
        h, w = (int(v) for v in rng.integers(1, hi, 2))
        yield (rng.random((h, w)) < rng.random()).astype(np.uint8)


def _check(img):
    h, w = img.shape
    b4 = O.canonical(O.bfs2d(img, 4)[0])[0]
    b8 = O.canonical(O.bfs2d(img, 8)[0])[0]
    r = O.runeq2d(img)
    assert np.array_equal(O.canonical(r)[0], b4)
    assert np.array_equal(r, V.runeq_values(img, (w + 1) // 2))
    rr = O.runeq2d(img, reference=True)
    assert np.array_equal(rr, V.runeq_values(img, O.load().orc_pad2(w) // 2))
    for conn, b in ((4, b4), (8, b8)):
        he = O.herun2d(img, conn)
        assert np.array_equal(O.canonical(he)[0], b)
        assert np.array_equal(he, V.herun_values(img, conn))
    g, n = O.grblock2d(img)
    assert np.array_equal(O.canonical(g)[0], b8) and n == int(b8.max())
    assert np.array_equal(g, V.block_order_values(img))


@pytest.mark.parametrize("shape", [(3, 4), (4, 3), (2, 6), (6, 2), (1, 9), (9, 1)])
def test_exhaustive(shape):
    for img in _all_images(*shape):
        _check(img)


def test_random():
    for img in _random_images(400, seed=5):
        _check(img)


def test_runeq_reference_slots():
    """The reference CPU numbers runs in w >> 1 slots per row (LabelingAlgs.cpp:561,
    578): for even widths that is ceil(w / 2), the default here."""
    for img in _random_images(100, seed=6):
        h, w = img.shape
        if w % 2 == 0:
            assert np.array_equal(O.runeq2d(img, slots=w >> 1), O.runeq2d(img))


def test_herun_quirks():
    """As written, He's labeler never scans the last row and closes runs one
    column early at the right edge (LabelingAlgs.cpp:150-171); with the last row
    and column empty the two restatements agree."""
    rng = np.random.default_rng(7)
    for _ in range(200):
        h, w = (int(v) for v in rng.integers(2, 16, 2))
        img = (rng.random((h, w)) < rng.random()).astype(np.uint8)
        img[-1, :] = 0
        img[:, -1] = 0
        for conn in (4, 8):
            assert np.array_equal(O.herun2d(img, conn, quirks=True), O.herun2d(img, conn, quirks=False))
    # known answer: a run reaching the last column loses that pixel, the last row is unlabelled
    img = np.array([[1, 1, 1], [0, 0, 0], [1, 0, 1]], np.uint8)
    assert O.herun2d(img, 4, quirks=True).tolist() == [[1, 1, 0], [0, 0, 0], [0, 0, 0]]
    assert O.herun2d(img, 4, quirks=False).tolist() == [[1, 1, 1], [0, 0, 0], [2, 0, 3]]


def test_value_examples():
    """Hand-worked answers.  A U shape: its two arms open labels 1 and 2, the
    base merges them and He keeps 1; the run below the U opens label 3 although
    only 2 components exist, so He's values skip numbers.  Run equivalence
    numbers the U by its left arm (row 0, run 0) and the bar by row 3, run 0;
    Grana's order follows first blocks."""
    img = np.array([[1, 0, 1, 0],
                    [1, 0, 1, 0],
                    [1, 1, 1, 0],
                    [0, 0, 0, 0],
                    [0, 1, 1, 1]], np.uint8)
    he = O.herun2d(img, 4)
    assert he[0, 0] == he[0, 2] == he[2, 1] == 1 and he[4, 1] == 3
    r = O.runeq2d(img)
    assert r[0, 0] == r[0, 2] == 1 and r[4, 1] == 1 + 4 * 2 + 0
    g, n = O.grblock2d(img)
    assert n == 2 and g[0, 0] == 1 and g[4, 3] == 2
    # diagonal contact: 8-connected for He's 8 mode and Grana, apart for the 4-connected methods
    d = np.array([[1, 0], [0, 1]], np.uint8)
    assert len(np.unique(O.herun2d(d, 8)[d > 0])) == 1 and len(np.unique(O.herun2d(d, 4)[d > 0])) == 2
    assert len(np.unique(O.runeq2d(d)[d > 0])) == 2 and O.grblock2d(d)[1] == 1
