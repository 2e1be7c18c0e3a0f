"""GPU parity of 3D labeling (bleq on 2x2x2 blocks and lbeq on voxels,
26-connectivity) against the oracle and the scipy goldens, through the C ABI.
Bit-exact: canonical labels and counts; raw labels = 1 + min block (bleq) or
min voxel (lbeq) index of each component in the unpadded [d][h][w] raster."""
import hashlib
import json
import os

import numpy as np
import pytest

import labeling_amd as L
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ALGS = ["bleq", "lbeq"]


def _small3d():
    z = np.load(os.path.join(GOLD, "small3d.npz"))
    names = sorted({k.split("__")[0] for k in z.files})
    return [(n, z[f"{n}__img"], z[f"{n}__l26"]) for n in names]


def _root_values(vol, canon, n, alg):
    d, h, w = vol.shape
    zz, yy, xx = np.nonzero(vol)
    if alg == "bleq":
        bh, bw = (h + 1) // 2, (w + 1) // 2
        key = ((zz // 2).astype(np.int64) * bh + yy // 2) * bw + xx // 2
    else:
        key = (zz.astype(np.int64) * h + yy) * w + xx
    comp = canon[zz, yy, xx]
    mins = np.full(n + 1, np.iinfo(np.int64).max)
    np.minimum.at(mins, comp, key)
    out = np.zeros(vol.shape, np.int64)
    out[zz, yy, xx] = mins[comp] + 1
    return out


@pytest.mark.parametrize("alg", ALGS)
@pytest.mark.parametrize("case", _small3d(), ids=lambda c: c[0])
def test_small_goldens(labeler, case, alg):
    name, vol, l26 = case
    print(f"[{alg}3d golden {name} {vol.shape}]", flush=True)
    lab, n, _ = labeler.label3d(vol, alg)
    c, m = O.canonical(lab)
    np.testing.assert_array_equal(c, l26)
    assert n == m == int(l26.max())


SHAPES = [(1, 1, 1), (2, 2, 2), (1, 1, 200), (200, 1, 1), (1, 200, 1), (3, 5, 7), (16, 16, 128), (17, 17, 129),
          (15, 9, 130), (33, 20, 70), (64, 64, 64), (40, 70, 260)]


@pytest.mark.parametrize("alg", ALGS)
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("p", [0.1, 0.3])
def test_vs_oracle_and_root_values(labeler, shape, p, alg):
    vol = L.bernoulli(shape, p, seed=41)
    print(f"[{alg}3d {shape} p={p}]", flush=True)
    lab, n, _ = labeler.label3d(vol, alg)
    expect = O.bleq3d(vol) if alg == "bleq" else O.lbeq3d(vol)
    e, m = O.canonical(expect)
    c, k = O.canonical(lab)
    np.testing.assert_array_equal(c, e)
    assert n == m == k
    canon, _ = O.bfs3d(vol, 26)
    np.testing.assert_array_equal(lab.astype(np.int64), _root_values(vol, canon, m, alg))


@pytest.mark.parametrize("alg", ALGS)
def test_blobs_mid(labeler, alg):
    vol = L.blobs3d(70, 90, 150, 3.0, 0.4, seed=5)
    lab, n, _ = labeler.label3d(vol, alg)
    canon, m = O.bfs3d(vol, 26)
    np.testing.assert_array_equal(O.canonical(lab)[0], canon)
    assert n == m


def test_3d_connectivity_errors(labeler):
    vol = np.zeros((4, 4, 4), np.uint8)
    for conn in (6, 18, 8):
        with pytest.raises(L.LabelingError) as e:
            labeler.label3d(vol, "bleq", conn=conn)
        assert e.value.status == L.EUNSUPPORTED
    lab, n, _ = labeler.label3d(vol, "bleq", conn=26)
    assert n == 0 and not lab.any()


def _large():
    path = os.path.join(GOLD, "large.json")
    return json.load(open(path)) if os.path.exists(path) else {}


FULL = [k for k in sorted(_large()) if k.startswith("c3_")]


@pytest.mark.slow
@pytest.mark.parametrize("alg", ALGS)
@pytest.mark.parametrize("name", FULL)
def test_full_size_vs_scipy_golden(labeler, name, alg):
    entry = _large()[name]
    spec = entry["spec"]
    print(f"[{alg}3d full {name}]", flush=True)
    if spec["kind"] == "bernoulli":
        vol = L.bernoulli(tuple(spec["shape"]), spec["p"], spec["seed"])
    else:
        vol = L.blobs3d(*spec["shape"], spec["sigma"], spec["q"], spec["seed"])
    assert hashlib.sha256(vol.tobytes()).hexdigest() == entry["input_sha256"]
    lab, n, _ = labeler.label3d(vol, alg)
    del vol
    c, m = O.canonical(lab)
    assert hashlib.sha256(c.tobytes()).hexdigest() == entry["conn26"]["labels_sha256"]
    assert n == m == entry["conn26"]["components"]


def test_counts_identical_over_repeated_runs(labeler):
    """Two hundred labelings of the same 512^3 Bernoulli p=0.1 volume (1.69 M
    components, the case with the most concurrent unions) from device memory:
    every run must report the same, correct component count (a lost union in
    the global merge would leave one extra root in some runs)."""
    from tests import hiprt

    entry = _large().get("c3_bern512_p1")
    if entry is None:
        pytest.skip("large golden manifest not present")
    spec, gold = entry["spec"], entry["conn26"]["components"]
    s = spec["shape"][0]
    vol = L.bernoulli(tuple(spec["shape"]), spec["p"], spec["seed"])
    d_v = hiprt.to_device(vol)
    d_lab = hiprt.DeviceBuffer(s * s * s * 4)
    d_n = hiprt.DeviceBuffer(4)
    del vol
    print(f"[3d repeated runs p=0.1, golden count {gold}]", flush=True)
    counts = set()
    for _ in range(200):
        labeler.label3d_dev(d_v.addr, s, s, s, d_lab.addr, "bleq", 0, d_n.addr)
        hiprt.synchronize()
        counts.add(int(d_n.download((1,), np.uint32)[0]))
    assert counts == {gold}, counts
