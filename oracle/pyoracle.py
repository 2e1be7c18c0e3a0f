"""ctypes binding of the CPU oracle (oracle/liblbl_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker or the timed CPU baseline.  The
product package never imports it.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblbl_oracle.so")
_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: run `make -C oracle`")
        lib = ctypes.CDLL(LIB_PATH)
        P, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
        pu32, pi, pd = ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
        sig = {
            "orc_pad2": (i64, [i64]),
            "orc_pad3": (i64, [i64]),
            "orc_bleq2d": (i32, [P, i64, i64, i64, P, i32, i32, pu32, pi, pd]),
            "orc_bleq3d": (i32, [P, i64, i64, i64, P, i32, pu32, pd]),
            "orc_lbeq2d": (i32, [P, i64, i64, i64, i32, P, i32, i32, pu32, pd]),
            "orc_lbeq3d": (i32, [P, i64, i64, i64, P, i32, pu32, pd]),
            "orc_bfs2d": (i64, [P, i64, i64, i64, i32, P]),
            "orc_bfs3d": (i64, [P, i64, i64, i64, i32, P]),
            "orc_canonicalize": (i64, [P, i64]),
            "orc_bleq2d_conn": (i64, [P, i64, i64, i64, P]),
            "orc_bleq3d_conn": (i64, [P, i64, i64, i64, P]),
            "orc_exhaustive2d": (i64, [i64, i64, i32, i32, ctypes.POINTER(ctypes.c_uint64)]),
            "orc_exhaustive3d": (i64, [i64, i64, i64, i32, ctypes.POINTER(ctypes.c_uint64)]),
            "orc_runeq2d": (i32, [P, i64, i64, i64, P, pu32]),
            "orc_herun2d": (i32, [P, i64, i64, i32, i32, P]),
            "orc_grblock2d": (i64, [P, i64, i64, P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _u8(a):
    return np.ascontiguousarray(a, dtype=np.uint8)


def bleq2d(img, padded=True, threads=0, with_info=False):
    """Restated TLabelEquivalenceX2 (LabelingAlgs.cpp:336-529); raw reference labels."""
    img = _u8(img)
    h, w = img.shape
    out = np.empty((h, w), dtype=np.uint32)
    it, oob, sec = ctypes.c_uint32(), ctypes.c_int(), ctypes.c_double()
    rc = load().orc_bleq2d(img.ctypes.data, h, w, w, out.ctypes.data, int(padded), threads, ctypes.byref(it),
                           ctypes.byref(oob), ctypes.byref(sec))
    if rc:
        raise ValueError(f"orc_bleq2d failed ({rc})")
    if with_info:
        return out, {"iterations": it.value, "oob": bool(oob.value), "seconds": sec.value}
    return out


def bleq3d(vol, threads=0, with_info=False):
    vol = _u8(vol)
    d, h, w = vol.shape
    out = np.empty((d, h, w), dtype=np.uint32)
    it, sec = ctypes.c_uint32(), ctypes.c_double()
    rc = load().orc_bleq3d(vol.ctypes.data, d, h, w, out.ctypes.data, threads, ctypes.byref(it), ctypes.byref(sec))
    if rc:
        raise ValueError(f"orc_bleq3d failed ({rc})")
    if with_info:
        return out, {"iterations": it.value, "seconds": sec.value}
    return out


def lbeq2d(img, conn=0, quirks=False, threads=0, with_info=False):
    img = _u8(img)
    h, w = img.shape
    out = np.empty((h, w), dtype=np.uint32)
    it, sec = ctypes.c_uint32(), ctypes.c_double()
    rc = load().orc_lbeq2d(img.ctypes.data, h, w, w, conn, out.ctypes.data, int(quirks), threads, ctypes.byref(it),
                           ctypes.byref(sec))
    if rc:
        raise ValueError(f"orc_lbeq2d failed ({rc})")
    if with_info:
        return out, {"iterations": it.value, "seconds": sec.value}
    return out


def lbeq3d(vol, threads=0, with_info=False):
    vol = _u8(vol)
    d, h, w = vol.shape
    out = np.empty((d, h, w), dtype=np.uint32)
    it, sec = ctypes.c_uint32(), ctypes.c_double()
    rc = load().orc_lbeq3d(vol.ctypes.data, d, h, w, out.ctypes.data, threads, ctypes.byref(it), ctypes.byref(sec))
    if rc:
        raise ValueError(f"orc_lbeq3d failed ({rc})")
    if with_info:
        return out, {"iterations": it.value, "seconds": sec.value}
    return out


def bfs2d(img, conn=8):
    img = _u8(img)
    h, w = img.shape
    out = np.empty((h, w), dtype=np.uint32)
    n = load().orc_bfs2d(img.ctypes.data, h, w, w, conn, out.ctypes.data)
    if n < 0:
        raise ValueError("orc_bfs2d failed")
    return out, int(n)


def bfs3d(vol, conn=26):
    vol = _u8(vol)
    d, h, w = vol.shape
    out = np.empty((d, h, w), dtype=np.uint32)
    n = load().orc_bfs3d(vol.ctypes.data, d, h, w, conn, out.ctypes.data)
    if n < 0:
        raise ValueError("orc_bfs3d failed")
    return out, int(n)


def canonical(labels):
    """Canonical relabelling (copy): 1..N in memory order of first appearance."""
    out = np.array(labels, dtype=np.uint32, copy=True, order="C")
    n = load().orc_canonicalize(out.ctypes.data, out.size)
    if n < 0:
        raise ValueError("orc_canonicalize failed")
    return out, int(n)


def bleq2d_conn(img):
    img = _u8(img)
    h, w = img.shape
    L = load()
    bw, bh = L.orc_pad2(w) // 2, L.orc_pad2(h) // 2
    out = np.empty((bh, bw), dtype=np.uint8)
    L.orc_bleq2d_conn(img.ctypes.data, h, w, w, out.ctypes.data)
    return out


def bleq3d_conn(vol):
    vol = _u8(vol)
    d, h, w = vol.shape
    L = load()
    shape = (L.orc_pad3(d) // 2, L.orc_pad3(h) // 2, L.orc_pad3(w) // 2)
    out = np.empty(shape, dtype=np.uint32)
    L.orc_bleq3d_conn(vol.ctypes.data, d, h, w, out.ctypes.data)
    return out


def exhaustive2d(h, w, alg, conn):
    """Run every h x w binary image through a restated labeler and the BFS.
    alg: 0 = bleq2d padded, 1 = lbeq2d.  Returns (mismatches, first failing image code)."""
    first = ctypes.c_uint64(0)
    n = load().orc_exhaustive2d(h, w, alg, conn, ctypes.byref(first))
    return int(n), int(first.value)


def exhaustive3d(d, h, w, alg):
    """Every d x h x w binary volume through bleq3d (alg 0) or lbeq3d (alg 1) vs BFS-26."""
    first = ctypes.c_uint64(0)
    n = load().orc_exhaustive3d(d, h, w, alg, ctypes.byref(first))
    return int(n), int(first.value)


# ---- the registry's other methods (oracle_registry.c) --------------------------------

def runeq2d(img, reference=False, slots=None):
    """Restated run equivalence (TRunEqivLabeling, LabelingAlgs.cpp:535-810).
    reference=False: the unpadded image with ceil(w/2) run slots per row (the
    CPU values, w >> 1 slots, for even w); reference=True: the reference GPU's
    padded image (orc_pad2 in both dimensions) with P2(w)/2 slots, cropped."""
    img = _u8(img)
    h, w = img.shape
    L = load()
    if reference:
        ph, pw = L.orc_pad2(h), L.orc_pad2(w)
        src = np.zeros((ph, pw), np.uint8)
        src[:h, :w] = img
        ns = pw // 2
    else:
        src, ph, pw = img, h, w
        ns = slots if slots is not None else (w + 1) // 2
    out = np.empty((ph, pw), dtype=np.uint32)
    it = ctypes.c_uint32()
    rc = L.orc_runeq2d(src.ctypes.data, ph, pw, ns, out.ctypes.data, ctypes.byref(it))
    if rc != 0:
        raise ValueError(f"orc_runeq2d failed ({rc})")
    return np.ascontiguousarray(out[:h, :w])


def herun2d(img, conn=4, quirks=False):
    """Restated He et al. run labeling (TRunLabeling, LabelingAlgs.cpp:58-205);
    conn 8 extends run contacts by one pixel (ConPix, :140)."""
    img = _u8(img)
    h, w = img.shape
    out = np.empty((h, w), dtype=np.uint32)
    rc = load().orc_herun2d(img.ctypes.data, h, w, 1 if conn == 8 else 0, int(quirks), out.ctypes.data)
    if rc != 0:
        raise ValueError("orc_herun2d failed")
    return out


def grblock2d(img):
    """Block-based two-pass 8-connected labeling after Grana et al.; (labels, N)."""
    img = _u8(img)
    h, w = img.shape
    out = np.empty((h, w), dtype=np.uint32)
    n = load().orc_grblock2d(img.ctypes.data, h, w, out.ctypes.data)
    if n < 0:
        raise ValueError("orc_grblock2d failed")
    return out, int(n)


def bin2d(img):
    """TBinLabeling (LabelingAlgs.cpp:23-36): 1 on foreground."""
    return (np.asarray(img) != 0).astype(np.uint32)
