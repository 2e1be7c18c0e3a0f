"""Continuous-batching server: every request's labels and component count in a
mixed, overlapping stream must equal the single-request result on the same
input (tolerance 0: labels are integers)."""
import numpy as np
import pytest

import labeling_amd as L
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def _mixed_requests(n, seed):
    reqs = []
    for i in range(n):
        u = L.synth_u64(seed, 0x5E, i)
        kind = u % 10
        if kind < 8:   # 2D
            h = 1 + L.synth_u64(seed, 0x5F, 2 * i) % 700
            w = 1 + L.synth_u64(seed, 0x5F, 2 * i + 1) % 700
            img = L.bernoulli((h, w), 0.1 + 0.8 * (u % 97) / 97, seed, stream=i)
            alg, conn = ("lbeq", 4 if kind == 7 else 8) if kind >= 6 else ("bleq", 0)
        else:          # 3D
            s = 2 + L.synth_u64(seed, 0x60, i) % 60
            img = L.bernoulli((s, s + 1, s + 3), 0.2 + 0.1 * (u % 3), seed, stream=i)
            alg, conn = ("bleq" if kind == 8 else "lbeq"), 0
        reqs.append((img, alg, conn))
    return reqs


def _run(labeler, reqs, **cfg):
    srv = L.Server(labeler, **cfg)
    pins = []
    tickets = {}
    try:
        for k, (img, alg, conn) in enumerate(reqs):
            pin_in = L.PinnedArray(img.shape, np.uint8)
            pin_in.array[...] = img
            pin_out = L.PinnedArray(img.shape, np.uint32)
            pins.append((pin_in, pin_out))
            t = srv.submit(pin_in.ptr, pin_out.ptr, img.shape, alg, conn)
            tickets[t] = k
        srv.drain()
        done = srv.poll(cap=len(reqs) + 8)
        assert len(done) == len(reqs)
        results = {}
        for ticket, status, ncomp, t_sub, t_start, t_done in done:
            assert status == L.OK
            assert t_sub <= t_start <= t_done
            k = tickets[ticket]
            results[k] = (pins[k][1].array.copy(), ncomp)
        return results
    finally:
        srv.close()
        for a, b in pins:
            a.free()
            b.free()


@pytest.mark.parametrize("cfg", [dict(streams=4, hbm_budget_bytes=256 << 20, max_batch=8),
                                 dict(streams=2, hbm_budget_bytes=12 << 20, max_batch=3),
                                 dict(streams=1, hbm_budget_bytes=64 << 20, max_batch=1)],
                         ids=["wide", "tight-memory", "serial"])
def test_mix_equals_alone(labeler, cfg):
    reqs = _mixed_requests(60, seed=4)
    print(f"[serve {cfg}]", flush=True)
    got = _run(labeler, reqs, **cfg)
    for k, (img, alg, conn) in enumerate(reqs):
        if img.ndim == 2:
            alone, n, _ = labeler.label2d(img, alg, conn=conn)
        else:
            alone, n, _ = labeler.label3d(img, alg)
        lab, ncomp = got[k]
        np.testing.assert_array_equal(lab, alone, err_msg=f"request {k} {alg} {img.shape}")
        assert ncomp == n


def test_admission_limits(labeler):
    srv = L.Server(labeler, streams=2, hbm_budget_bytes=2 << 20, max_batch=4)
    try:
        big = L.PinnedArray((1024, 1024), np.uint8)
        out = L.PinnedArray((1024, 1024), np.uint32)
        with pytest.raises(L.LabelingError) as e:   # 1 MiB in + 4 MiB labels > 2 MiB arena
            srv.submit(big.ptr, out.ptr, (1024, 1024), "bleq")
        assert e.value.status == L.ENOMEM
        with pytest.raises(L.LabelingError) as e:
            srv.submit(big.ptr, out.ptr, (16, 16), "bleq", conn=4)
        assert e.value.status == L.EUNSUPPORTED
        big.free()
        out.free()
    finally:
        srv.close()


def test_canonical_mode_mix_equals_alone(labeler):
    """In LBL_LABELS_CANONICAL mode the server labels every request with the
    device canonical numbering (2D bleq requests are not fused then)."""
    reqs = _mixed_requests(30, seed=9)
    labeler.set_mode(L.LABELS_CANONICAL)
    try:
        print("[serve canonical]", flush=True)
        got = _run(labeler, reqs, streams=3, hbm_budget_bytes=64 << 20, max_batch=4)
        for k, (img, alg, conn) in enumerate(reqs):
            if img.ndim == 2:
                alone, n, _ = labeler.label2d(img, alg, conn=conn)
            else:
                alone, n, _ = labeler.label3d(img, alg)
            lab, ncomp = got[k]
            np.testing.assert_array_equal(lab, alone, err_msg=f"request {k} {alg} {img.shape}")
            assert ncomp == n
            assert lab.max() == n   # canonical: labels 1..n
    finally:
        labeler.set_mode(L.LABELS_ROOT)


@pytest.mark.timeout(900)
def test_serving_workload_vs_oracle(labeler):
    """Requests of the serving configuration itself (2D sides 256..4096, 3D
    sides 64..256, half blobs, the bench's generator), submitted together to a
    server with 3 streams so that they overlap: each result equals the
    oracle's labeling (canonical labels and count, the independent BFS) and
    the single-request result on the same input."""
    n = 36
    srv = L.Server(labeler, streams=3, hbm_budget_bytes=3 << 30, max_batch=8)
    reqs, tickets = [], {}
    try:
        for i in range(n):
            img = L.serving_request(i, seed=11)
            alg = "lbeq" if i % 6 == 5 else "bleq"
            conn = 8 if alg == "lbeq" and img.ndim == 2 else 0
            pin_in = L.PinnedArray(img.shape, np.uint8)
            pin_in.array[...] = img
            pin_out = L.PinnedArray(img.shape, np.uint32)
            reqs.append((img, alg, conn, pin_in, pin_out))
        print(f"[serve C5] {n} requests, sides {[r[0].shape[0] for r in reqs]}", flush=True)
        for k, (img, alg, conn, pin_in, pin_out) in enumerate(reqs):
            tickets[srv.submit(pin_in.ptr, pin_out.ptr, img.shape, alg, conn)] = k
        srv.drain()
        done = srv.poll(cap=n + 8)
        assert len(done) == n
        for ticket, status, ncomp, t_sub, t_start, t_done in done:
            k = tickets[ticket]
            img, alg, conn, _, pin_out = reqs[k]
            assert status == L.OK, f"request {k}"
            lab = pin_out.array
            expect, m = O.bfs2d(img, conn or 8) if img.ndim == 2 else O.bfs3d(img, 26)
            np.testing.assert_array_equal(O.canonical(lab)[0], expect, err_msg=f"request {k} {alg} {img.shape}")
            assert ncomp == m, f"request {k}"
            if img.ndim == 2:
                alone, n1, _ = labeler.label2d(img, alg, conn=conn)
            else:
                alone, n1, _ = labeler.label3d(img, alg)
            np.testing.assert_array_equal(lab, alone, err_msg=f"request {k} alone vs mix")
            assert n1 == ncomp
    finally:
        srv.close()
        for r in reqs:
            r[3].free()
            r[4].free()


def test_injected_launch_fault_surfaces(labeler):
    """A group whose kernel launch fails completes with LBL_EDEVICE for each of
    its requests (labels not written); the server keeps serving."""
    imgs = [L.bernoulli((200 + 10 * i, 300), 0.5, seed=i) for i in range(6)]
    srv = L.Server(labeler, streams=1, hbm_budget_bytes=64 << 20, max_batch=3)
    pins = []
    try:
        srv.inject_faults(1)
        tickets = {}
        for k, img in enumerate(imgs):
            pi, po = L.PinnedArray(img.shape, np.uint8), L.PinnedArray(img.shape, np.uint32)
            pi.array[...] = img
            po.array[...] = 7
            pins.append((pi, po))
            tickets[srv.submit(pi.ptr, po.ptr, img.shape, "bleq")] = k
        srv.drain()
        done = sorted(srv.poll(cap=16), key=lambda c: c[0])
        statuses = [c[1] for c in done]
        assert statuses[0] == L.EDEVICE
        assert L.OK in statuses and set(statuses) <= {L.OK, L.EDEVICE}
        for ticket, status, ncomp, *_ in done:
            k = tickets[ticket]
            if status == L.EDEVICE:
                assert (pins[k][1].array == 7).all()   # untouched
            else:
                alone, n, _ = labeler.label2d(imgs[k], "bleq")
                np.testing.assert_array_equal(pins[k][1].array, alone)
                assert ncomp == n
    finally:
        srv.close()
        for a, b in pins:
            a.free()
            b.free()
