"""GPU parity of 2D block equivalence (bleq) against the oracle, through the C ABI.

Bar: bit-exact.  Canonical labels and component counts must equal the oracle's
(and the scipy goldens); in LBL_LABELS_REFERENCE mode the raw label values
must equal the reference OpenCL values restated by the oracle
(1 + min padded-block index, LabelingAlgs.cl:202,317).
"""
import ctypes
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
    return [(n, z[f"{n}__img"], z[f"{n}__l8"]) for n in names]


@pytest.fixture
def ref_mode(labeler):
    labeler.set_mode(L.LABELS_REFERENCE)
    yield labeler
    labeler.set_mode(L.LABELS_ROOT)


@pytest.mark.parametrize("case", _small2d(), ids=lambda c: c[0])
def test_small_goldens(labeler, case):
    name, img, l8 = case
    print(f"[bleq2d golden {name} {img.shape}]", flush=True)
    lab, n, _ = labeler.label2d(img, "bleq")
    c, m = O.canonical(lab)
    np.testing.assert_array_equal(c, l8)
    assert n == m == int(l8.max())


SHAPES = [(1, 1), (2, 2), (1, 300), (300, 1), (3, 5), (31, 257), (32, 256), (33, 255), (34, 258), (64, 512),
          (65, 513), (257, 511), (100, 1000), (1000, 100), (517, 1029), (1024, 1024)]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"{s[0]}x{s[1]}")
@pytest.mark.parametrize("p", [0.2, 0.5, 0.8])
def test_reference_raw_values(ref_mode, shape, p):
    img = L.bernoulli(shape, p, seed=hash((shape, p)) & 0xFFFF)
    print(f"[bleq2d raw {shape} p={p}]", flush=True)
    lab, n, _ = ref_mode.label2d(img, "bleq")
    expect = O.bleq2d(img)            # padded reference restatement: raw values
    np.testing.assert_array_equal(lab, expect)
    assert n == O.canonical(expect)[1]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"{s[0]}x{s[1]}")
def test_root_labels_vs_bfs(labeler, shape):
    img = L.bernoulli(shape, 0.55, seed=17)
    print(f"[bleq2d root {shape}]", flush=True)
    lab, n, _ = labeler.label2d(img, "bleq")
    canon, m = O.bfs2d(img, 8)
    c, k = O.canonical(lab)
    np.testing.assert_array_equal(c, canon)
    assert n == m == k
    # root labels: 1 + min index in the unpadded block grid
    bw = (shape[1] + 1) // 2
    ys, xs = np.nonzero(img)
    if len(ys):
        blk = (ys // 2) * bw + xs // 2
        comp = canon[ys, xs]
        mins = np.full(m + 1, np.iinfo(np.int64).max)
        np.minimum.at(mins, comp, blk)
        np.testing.assert_array_equal(lab[ys, xs], mins[comp] + 1)


def test_blobs_and_serpentine(labeler):
    img = L.blobs2d(777, 1555, 4.0, 0.5, seed=3)
    lab, n, _ = labeler.label2d(img, "bleq")
    np.testing.assert_array_equal(O.canonical(lab)[0], O.bfs2d(img, 8)[0])
    ser = np.zeros((301, 999), np.uint8)
    for r in range(0, 301, 2):
        ser[r, :] = 255
        if r + 1 < 301:
            ser[r + 1, 998 if (r // 2) % 2 == 0 else 0] = 255
    lab, n, _ = labeler.label2d(ser, "bleq")
    assert n == 1 and O.canonical(lab)[1] == 1


def test_deterministic(labeler):
    img = L.bernoulli((1500, 2500), 0.5, seed=5)
    a = labeler.label2d(img, "bleq")[0]
    for _ in range(3):
        np.testing.assert_array_equal(labeler.label2d(img, "bleq")[0], a)


def test_host_pitch(labeler):
    """Rows with a pitch larger than the width (host entry point)."""
    h, w, pitch = 123, 301, 333
    buf = L.bernoulli((h, pitch), 0.5, seed=9)
    out = np.empty((h, w), np.uint32)
    n = ctypes.c_uint32()
    st = labeler.lib.lbl_label2d(labeler.handle, L.BLEQ, buf.ctypes.data, h, w, pitch, 0, out.ctypes.data,
                                 ctypes.byref(n), None)
    assert st == L.OK
    img = np.ascontiguousarray(buf[:, :w])
    np.testing.assert_array_equal(O.canonical(out)[0], O.bfs2d(img, 8)[0])


def test_device_entry_point_unaligned_pitch(labeler):
    """Device buffers with a pitch that is not a multiple of 16 (scalar load path)."""
    from tests import hiprt

    h, w, pitch = 257, 1001, 1003
    host = L.bernoulli((h, pitch), 0.45, seed=21)
    d_px = hiprt.to_device(host)
    d_lab = hiprt.DeviceBuffer(h * w * 4 + 4)
    d_n = hiprt.DeviceBuffer(4)
    stream = hiprt.Stream()
    # labels at a 4-byte (not 16-byte) aligned address: scalar store path
    labeler.label2d_dev(d_px.addr, h, w, pitch, d_lab.addr + 4, "bleq", 0, d_n.addr, stream.handle.value)
    stream.synchronize()
    lab = d_lab.download((h, w), np.uint32, offset=4)
    canon, m = O.bfs2d(np.ascontiguousarray(host[:, :w]), 8)
    np.testing.assert_array_equal(O.canonical(lab)[0], canon)
    assert int(d_n.download((1,), np.uint32)[0]) == m
    stream.destroy()


def test_errors(labeler):
    img = np.zeros((4, 4), np.uint8)
    with pytest.raises(L.LabelingError) as e:
        labeler.label2d(img, "bleq", conn=4)
    assert e.value.status == L.EUNSUPPORTED
    with pytest.raises(L.LabelingError) as e:
        labeler.label2d(img, "bleq", conn=6)
    assert e.value.status == L.EINVAL
    with pytest.raises(L.LabelingError) as e:
        labeler.label2d(np.zeros((0, 4), np.uint8), "bleq")
    assert e.value.status == L.EINVAL


def _large():
    path = os.path.join(GOLD, "large.json")
    return json.load(open(path)) if os.path.exists(path) else {}


FULL = [k for k in sorted(_large()) if k.startswith("c2_") or k.startswith("c1_")]


@pytest.mark.slow
@pytest.mark.parametrize("name", FULL)
def test_full_size_vs_scipy_golden(labeler, name):
    entry = _large()[name]
    spec = entry["spec"]
    print(f"[bleq2d full {name}]", flush=True)
    if spec["kind"] == "bernoulli":
        img = L.bernoulli(tuple(spec["shape"]), spec["p"], spec["seed"])
    else:
        img = L.blobs2d(*spec["shape"], spec["sigma"], spec["q"], spec["seed"])
    assert hashlib.sha256(img.tobytes()).hexdigest() == entry["input_sha256"]
    lab, n, _ = labeler.label2d(img, "bleq")
    c, m = O.canonical(lab)
    assert hashlib.sha256(c.tobytes()).hexdigest() == entry["conn8"]["labels_sha256"]
    assert n == m == entry["conn8"]["components"]


@pytest.mark.parametrize("flatten", [1, 0])
@pytest.mark.parametrize("tile", range(7))
def test_every_tile_shape_and_schedule(tmp_path, tile, flatten):
    """Each launch configuration the autotuner can pick gives the same labels."""
    import math

    lab = L.Labeler(0)   # own context: its tuning table forces the configuration
    try:
        for k, (h, w) in enumerate([(517, 1029), (1024, 1024), (33, 255), (300, 2047)]):
            cls = int(math.floor(math.log2(h * w)))
            path = tmp_path / f"t{k}.json"
            path.write_text(json.dumps({f"bleq/single2d/{cls}": [tile, flatten, 0.0]}))
            assert lab.lib.lbl_tuning_load(lab.handle, str(path).encode()) == L.OK
            img = L.bernoulli((h, w), 0.3 + 0.1 * k, seed=50 + k)
            got, n, _ = lab.label2d(img, "bleq")
            canon, m = O.bfs2d(img, 8)
            np.testing.assert_array_equal(O.canonical(got)[0], canon, err_msg=f"tile {tile} {h}x{w}")
            assert n == m
            assert lab.tuning()[f"bleq/single2d/{cls}"][:2] == [tile, flatten]
    finally:
        lab.close()


def test_autotune_then_label(tmp_path):
    from tests import hiprt

    lab = L.Labeler(0)
    try:
        h, w = 1500, 2300
        img = L.bernoulli((h, w), 0.5, seed=77)
        d_px = hiprt.to_device(img)
        d_lab = hiprt.DeviceBuffer(h * w * 4)
        lab.autotune2d(d_px.addr, h, w, w, d_lab.addr)
        table = lab.tuning()
        assert any(k.startswith("bleq/single2d/") for k in table), table
        got, n, _ = lab.label2d(img, "bleq")
        canon, m = O.bfs2d(img, 8)
        np.testing.assert_array_equal(O.canonical(got)[0], canon)
        p = tmp_path / "tuning.json"
        assert lab.lib.lbl_tuning_save(lab.handle, str(p).encode()) == L.OK
        assert json.loads(p.read_text()) == table
    finally:
        lab.close()


def test_autotune_orders_after_the_callers_stream():
    """lbl_autotune2d(stream=S): the timed launches run after what S holds at
    the call (here the image, copied in on S with no host wait), and S waits
    for them on return (the labels they write, copied out on S)."""
    from tests import hiprt

    lab = L.Labeler(0)
    s = hiprt.Stream()
    try:
        h, w = 2048, 2048
        img = L.bernoulli((h, w), 0.5, seed=5)
        d_src = hiprt.to_device(img)
        d_px = hiprt.DeviceBuffer(h * w)
        d_px.zero()
        d_lab = hiprt.DeviceBuffer(h * w * 4)
        d_out = hiprt.DeviceBuffer(h * w * 4)
        d_out.zero()
        hiprt.copy_async(d_px.addr, d_src.addr, h * w, s)
        lab.autotune2d(d_px.addr, h, w, w, d_lab.addr, stream=s.handle.value)
        hiprt.copy_async(d_out.addr, d_lab.addr, h * w * 4, s)
        s.synchronize()
        got = d_out.download((h, w), np.int32)
        canon, m = O.bfs2d(img, 8)
        np.testing.assert_array_equal(O.canonical(got)[0], canon)
    finally:
        s.destroy()
        lab.close()


def test_counts_identical_over_repeated_runs(labeler):
    """One hundred labelings of the 8192^2 Bernoulli p=0.5 golden input (one
    component spanning the image plus 220 k others: the heaviest union work
    in the tile and merge kernels) from device memory: every run must report
    the golden component count."""
    from tests import hiprt

    entry = _large().get("c2_bern8192_p5")
    if entry is None:
        pytest.skip("large golden manifest not present")
    spec, gold = entry["spec"], entry["conn8"]["components"]
    h, w = spec["shape"]
    img = L.bernoulli((h, w), spec["p"], spec["seed"])
    d_px = hiprt.to_device(img)
    d_lab = hiprt.DeviceBuffer(h * w * 4)
    d_n = hiprt.DeviceBuffer(4)
    del img
    print(f"[bleq2d repeated runs p=0.5, golden count {gold}]", flush=True)
    counts = set()
    for _ in range(100):
        labeler.label2d_dev(d_px.addr, h, w, w, d_lab.addr, "bleq", 0, d_n.addr)
        hiprt.synchronize()
        counts.add(int(d_n.download((1,), np.uint32)[0]))
    assert counts == {gold}, counts
