"""Closed forms of the label values of the reference registry's other methods,
computed with numpy from a component labeling (BFS) and the image's runs, by
rules that follow from each algorithm's definition, not from its code:

* run equivalence (TRunEqivLabeling, LabelingAlgs.cpp:535-810): runs are
  numbered 1 + row * slots + (index in the row); a component takes the
  smallest number among its runs, which is its first run's (topmost row, then
  leftmost): the run holding the component's first pixel.
* He et al.'s run labeling (TRunLabeling, LabelingAlgs.cpp:58-205): a run opens
  a new provisional label (1, 2, 3, ... in raster order) exactly when no run of
  the row above lies in reach (columns start - conpix .. end + conpix); merges
  keep the smaller label, so a component keeps the label its first run opened
  (a first run never has a run of its component above it).
* Grana et al.'s block labeling (cvlabeling_imagelab.cpp): provisional labels
  per 2x2 block in block raster order, merges keep the smaller, final labels
  1..N in order of the representatives: components ordered by their first
  block (the minimum block index).

Used by the tests to pin the oracle's restatements and, at sizes the oracle
cannot finish quickly, to check the GPU directly.
"""
import numpy as np


def _components(img, conn):
    from oracle import pyoracle as O

    lab, n = O.bfs2d(img, conn)
    return lab, n


def _run_starts(img):
    """Boolean mask of run starts (a foreground pixel whose left neighbour is not)."""
    left = np.zeros_like(img)
    left[:, 1:] = img[:, :-1]
    return img & ~left


def first_pixels(lab, n):
    """(y, x) of every component's first pixel in raster order, by component id 1..n."""
    flat = lab.reshape(-1)
    idx = np.nonzero(flat)[0]
    ids, pos = np.unique(flat[idx], return_index=True)
    first = np.full(n + 1, -1, np.int64)
    first[ids] = idx[pos]
    w = lab.shape[1]
    return first // w, first % w


def runeq_values(img, slots):
    """1 + y * slots + k for the component's first run, the k-th of row y (its
    first pixel starts a run: the pixel to its left is not in the image's
    foreground, or it would be an earlier pixel of the component)."""
    img = np.asarray(img) != 0
    lab, n = _components(img.astype(np.uint8), 4)
    fy, fx = first_pixels(lab, n)
    st = _run_starts(img).astype(np.int64)
    before = np.cumsum(st, axis=1) - st   # run starts of the row left of each pixel
    value = np.zeros(n + 1, np.int64)
    value[1:] = 1 + fy[1:] * slots + before[fy[1:], fx[1:]]
    return value[lab].astype(np.uint32) * (lab > 0)


def herun_values(img, conn):
    """The provisional label of each component's first run (see above)."""
    img = np.asarray(img) != 0
    h, w = img.shape
    conpix = 1 if conn == 8 else 0
    lab, n = _components(img.astype(np.uint8), conn)
    st = _run_starts(img)
    right = np.zeros_like(img)
    right[:, :-1] = img[:, 1:]
    ys, ss = np.nonzero(st)                          # raster order
    ye, ee = np.nonzero(img & ~right)                # run ends, same order
    assert np.array_equal(ys, ye)
    # pixels of the row above in columns s - conpix .. e + conpix (prefix sums per row)
    up = np.zeros((h, w + 1), np.int64)
    up[1:, 1:] = np.cumsum(img[:-1].astype(np.int64), axis=1)
    lo = np.maximum(ss - conpix, 0)
    hi = np.minimum(ee + conpix, w - 1)
    opens = (up[ys, hi + 1] - up[ys, lo]) == 0
    rank = np.cumsum(opens)                          # provisional label of each opening run
    runrank = np.zeros((h, w), np.int64)
    runrank[ys, ss] = np.where(opens, rank, 0)
    fy, fx = first_pixels(lab, n)
    value = np.zeros(n + 1, np.int64)
    value[1:] = runrank[fy[1:], fx[1:]]
    assert (value[1:] > 0).all()                     # every first run opens a label
    return value[lab].astype(np.uint32) * (lab > 0)


def block_order_values(img):
    """1..N by each component's first 2x2 block in block raster order (8-connectivity)."""
    img = np.asarray(img) != 0
    h, w = img.shape
    lab, n = _components(img.astype(np.uint8), 8)
    ys, xs = np.nonzero(lab)
    bw = (w + 1) // 2
    blocks = (ys // 2) * bw + xs // 2
    first = np.full(n + 1, np.iinfo(np.int64).max)
    np.minimum.at(first, lab[ys, xs], blocks)
    order = np.argsort(first[1:], kind="stable")
    value = np.zeros(n + 1, np.int64)
    value[order + 1] = np.arange(1, n + 1)
    return value[lab].astype(np.uint32) * (lab > 0)
