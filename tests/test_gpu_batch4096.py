"""The batched configuration itself: 4096 images with H, W ~ U{64..2048} and
p ~ U[0.05, 0.95] (L.batched_workload(4096, 0)), labelled by one autotuned,
graph-captured batch exactly as bench.py runs it, checked image by image
against scipy goldens (tests/golden/batch4096.json: component count and the
SHA-256 of the canonical 8-connected labels).  The batch replaces the
reference's sequential per-image loop (ProcessImages, main.cpp:309-340)."""
import ctypes
import hashlib
import json
import os

import numpy as np
import pytest

import labeling_amd as L
from oracle import pyoracle as O
from tests import hiprt

pytestmark = [pytest.mark.gpu, pytest.mark.slow]
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "batch4096.json")


@pytest.fixture(scope="module")
def workload():
    gold = json.load(open(GOLD))
    specs = L.batched_workload(gold["spec"]["n"], gold["spec"]["seed"])
    assert [(h, w) for h, w, _ in specs] == [(g[0], g[1]) for g in gold["images"]]
    in_off, lab_off, o_in, o_lab = [], [], 0, 0
    for h, w, _ in specs:   # bench.py's layout: 16-byte pitched rows, dense labels
        in_off.append(o_in)
        lab_off.append(o_lab)
        o_in += h * ((w + 15) // 16 * 16)
        o_lab += (h * w + 3) // 4 * 4
    d_in = hiprt.DeviceBuffer(o_in)
    d_lab = hiprt.DeviceBuffer(4 * o_lab)
    for i, (h, w, p) in enumerate(specs):
        pitch = (w + 15) // 16 * 16
        img = np.zeros((h, pitch), np.uint8)
        img[:, :w] = L.bernoulli((h, w), p, seed=gold["spec"]["seed"], stream=i)
        d_in.upload(img, in_off[i])
    descs = (L.ImageDesc * len(specs))()
    for i, (h, w, _) in enumerate(specs):
        descs[i] = L.ImageDesc(d_in.addr + in_off[i], d_lab.addr + 4 * lab_off[i], h, w, (w + 15) // 16 * 16)
    print(f"[batch4096] {len(specs)} images, {gold['pixels'] / 1e9:.3f} Gpx uploaded", flush=True)
    yield gold, specs, descs, d_lab, lab_off
    d_in.free()
    d_lab.free()


def _check(gold, specs, d_lab, lab_off, counts, canonical_on_device, every=1):
    bad = []
    for i in range(0, len(specs), every):
        h, w, _ = specs[i]
        g = gold["images"][i]
        lab = d_lab.download((h, w), np.uint32, 4 * lab_off[i])
        c, m = (lab, int(lab.max())) if canonical_on_device else O.canonical(lab)
        if counts[i] != g[2] or m != g[2] or hashlib.sha256(c.tobytes()).hexdigest() != g[3]:
            bad.append((i, h, w, int(counts[i]), m, g[2]))
    assert not bad, f"{len(bad)} images differ, first: {bad[:5]}"


@pytest.mark.timeout(900)
def test_batch4096_autotuned_graph_vs_goldens(labeler, workload):
    gold, specs, descs, d_lab, lab_off = workload
    lib = labeler.lib
    n = len(specs)
    d_n = hiprt.DeviceBuffer(4 * n)
    b = ctypes.c_void_p()
    st = lib.lbl_batch_create(labeler.handle, L.BLEQ, descs, n, 0, ctypes.c_void_p(d_n.addr), L.BATCH_AUTOTUNE,
                              ctypes.byref(b))
    assert st == L.OK, lib.lbl_last_error(labeler.handle)
    try:
        t, tw, th, fl = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_int()
        lib.lbl_batch_config(b, ctypes.byref(t), ctypes.byref(tw), ctypes.byref(th), ctypes.byref(fl))
        print(f"[batch4096] tile {t.value} ({tw.value}x{th.value} blocks) flatten={fl.value}", flush=True)
        for rep in range(2):   # graph replays, with eager work between them
            d_lab.zero()
            d_n.zero()
            assert lib.lbl_batch_run(b, None) == L.OK
            hiprt.synchronize()
            counts = d_n.download((n,), np.uint32)
            _check(gold, specs, d_lab, lab_off, counts, False, every=1 if rep == 0 else 7)
            print(f"[batch4096] replay {rep} matches the goldens", flush=True)
    finally:
        lib.lbl_batch_destroy(b)
        d_n.free()


@pytest.mark.timeout(900)
def test_batch4096_canonical_on_device(labeler, workload):
    """The same batch in LBL_LABELS_CANONICAL mode: every image's labels are
    its canonical numbering, computed on the device."""
    gold, specs, descs, d_lab, lab_off = workload
    lib = labeler.lib
    n = len(specs)
    d_n = hiprt.DeviceBuffer(4 * n)
    b = ctypes.c_void_p()
    labeler.set_mode(L.LABELS_CANONICAL)
    try:
        st = lib.lbl_batch_create(labeler.handle, L.BLEQ, descs, n, 0, ctypes.c_void_p(d_n.addr), 0,
                                  ctypes.byref(b))
        assert st == L.OK, lib.lbl_last_error(labeler.handle)
        d_lab.zero()
        assert lib.lbl_batch_run(b, None) == L.OK
        hiprt.synchronize()
        counts = d_n.download((n,), np.uint32)
        _check(gold, specs, d_lab, lab_off, counts, True, every=3)
    finally:
        labeler.set_mode(L.LABELS_ROOT)
        if b.value:
            lib.lbl_batch_destroy(b)
        d_n.free()
