"""MI355X-native connected-component labeling (block equivalence and label
equivalence) behind the C ABI declared in include/labeling_amd.h.

This module is a thin ctypes binding used by the tests and the benchmark; the
product is the native library ``liblabeling_amd.so`` (HIP kernels for gfx950 +
C++ host code).  Loading fails loudly when the library has not been built: there
is no CPU fallback.

Reference interface mirrored: ``ILabeling::Label(pixels, labels, threads, coh)``
(LabelingTools.hpp:64,110,154) with the algorithm chosen by name from
``ALG_LIST`` (main.cpp:97-131, SetLabelingAlg main.cpp:373-395).
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblabeling_amd.so")
SYNTH_PATH = os.path.join(_HERE, "libsynth.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "labeling_amd.h")

# lbl_status
OK, EINVAL, EUNSUPPORTED, ENOMEM, EDEVICE, EITER, EBUSY = range(7)
STATUS_NAMES = ["LBL_OK", "LBL_EINVAL", "LBL_EUNSUPPORTED", "LBL_ENOMEM", "LBL_EDEVICE", "LBL_EITER", "LBL_EBUSY"]
# lbl_alg
BLEQ, LBEQ = 0, 1
ALGS = {"bleq": BLEQ, "lbeq": LBEQ}
# lbl_label_mode
LABELS_ROOT, LABELS_REFERENCE, LABELS_CANONICAL = 0, 1, 2

_lib = None
_synth = None


class LabelingError(RuntimeError):
    """A non-OK lbl_status, with the context's last error message."""

    def __init__(self, status: int, message: str):
        self.status = status
        name = STATUS_NAMES[status] if 0 <= status < len(STATUS_NAMES) else str(status)
        super().__init__(f"{name}: {message}")


def _declare(lib):
    P = ctypes.c_void_p
    u32, i32, sz = ctypes.c_uint32, ctypes.c_int, ctypes.c_size_t
    pu32 = ctypes.POINTER(ctypes.c_uint32)
    pf = ctypes.POINTER(ctypes.c_float)
    sig = {
        "lbl_create": (i32, [ctypes.POINTER(P), i32]),
        "lbl_destroy": (None, [P]),
        "lbl_last_error": (ctypes.c_char_p, [P]),
        "lbl_status_string": (ctypes.c_char_p, [i32]),
        "lbl_set_label_mode": (i32, [P, i32]),
        "lbl_set_iteration_cap": (i32, [P, u32]),
        "lbl_alg_from_name": (i32, [ctypes.c_char_p, ctypes.POINTER(i32)]),
        "lbl_label2d": (i32, [P, i32, P, u32, u32, sz, i32, P, pu32, pf]),
        "lbl_label3d": (i32, [P, i32, P, u32, u32, u32, i32, P, pu32, pf]),
        "lbl_label2d_dev": (i32, [P, i32, P, u32, u32, sz, i32, P, P, P]),
        "lbl_label3d_dev": (i32, [P, i32, P, u32, u32, u32, i32, P, P, P]),
        "lbl_set_profiling": (i32, [P, i32]),
        "lbl_last_kernel_times": (i32, [P, pf, ctypes.POINTER(ctypes.c_char_p), i32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def load() -> ctypes.CDLL:
    """Load liblabeling_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: run `make -C labeling_amd` (or __graft_entry__.build())")
        _lib = _declare(ctypes.CDLL(LIB_PATH))
    return _lib


def load_synth() -> ctypes.CDLL:
    global _synth
    if _synth is None:
        if not os.path.exists(SYNTH_PATH):
            raise OSError(f"{SYNTH_PATH} is missing: run `make -C labeling_amd`")
        lib = ctypes.CDLL(SYNTH_PATH)
        lib.lbl_synth_mix64.restype = ctypes.c_uint64
        lib.lbl_synth_mix64.argtypes = [ctypes.c_uint64]
        lib.lbl_synth_u64.restype = ctypes.c_uint64
        lib.lbl_synth_u64.argtypes = [ctypes.c_uint64] * 3
        lib.lbl_synth_bernoulli.restype = None
        lib.lbl_synth_bernoulli.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_uint64,
                                            ctypes.c_uint64]
        lib.lbl_synth_blobs2d.restype = ctypes.c_int
        lib.lbl_synth_blobs2d.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_uint64]
        lib.lbl_synth_blobs3d.restype = ctypes.c_int
        lib.lbl_synth_blobs3d.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_uint64]
        _synth = lib
    return _synth


# ---- synthetic inputs ---------------------------------------------------------

def bernoulli(shape, p: float, seed: int, stream: int = 0) -> np.ndarray:
    """uint8 array of `shape`, 255 with probability p (counter-based, portable)."""
    out = np.empty(shape, dtype=np.uint8)
    load_synth().lbl_synth_bernoulli(out.ctypes.data, out.size, float(p), seed, stream)
    return out


def blobs2d(h: int, w: int, sigma: float, quantile: float, seed: int) -> np.ndarray:
    out = np.empty((h, w), dtype=np.uint8)
    if load_synth().lbl_synth_blobs2d(out.ctypes.data, h, w, sigma, quantile, seed) != 0:
        raise ValueError("bad blob parameters")
    return out


def blobs3d(d: int, h: int, w: int, sigma: float, quantile: float, seed: int) -> np.ndarray:
    out = np.empty((d, h, w), dtype=np.uint8)
    if load_synth().lbl_synth_blobs3d(out.ctypes.data, d, h, w, sigma, quantile, seed) != 0:
        raise ValueError("bad blob parameters")
    return out


def synth_u64(seed: int, stream: int, i: int) -> int:
    return int(load_synth().lbl_synth_u64(seed, stream, i))


# ---- labeling ---------------------------------------------------------------------

def _alg_id(alg) -> int:
    if isinstance(alg, str):
        if alg not in ALGS:
            raise LabelingError(EUNSUPPORTED if alg in ("bin", "runeq", "he-run", "gr-block", "ocv") else EINVAL,
                                f"algorithm {alg!r} is not provided")
        return ALGS[alg]
    return int(alg)


class Labeler:
    """One lbl_ctx: a device, a stream and cached device buffers."""

    def __init__(self, device: int = 0):
        self.lib = load()
        self._ctx = ctypes.c_void_p()
        st = self.lib.lbl_create(ctypes.byref(self._ctx), device)
        if st != OK:
            raise LabelingError(st, self.lib.lbl_last_error(None).decode())

    def close(self):
        if self._ctx:
            self.lib.lbl_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._ctx

    def _check(self, st: int):
        if st != OK:
            raise LabelingError(st, self.lib.lbl_last_error(self._ctx).decode())

    def set_mode(self, mode: int):
        self._check(self.lib.lbl_set_label_mode(self._ctx, mode))

    def set_iteration_cap(self, cap: int):
        self._check(self.lib.lbl_set_iteration_cap(self._ctx, cap))

    def set_profiling(self, on: bool):
        self._check(self.lib.lbl_set_profiling(self._ctx, 1 if on else 0))

    def kernel_times(self):
        us = (ctypes.c_float * 16)()
        names = (ctypes.c_char_p * 16)()
        n = self.lib.lbl_last_kernel_times(self._ctx, us, names, 16)
        return [(names[i].decode(), float(us[i])) for i in range(max(n, 0))]

    def label2d(self, img: np.ndarray, alg="bleq", conn: int = 0) -> Tuple[np.ndarray, int, float]:
        img = np.ascontiguousarray(img, dtype=np.uint8)
        if img.ndim != 2:
            raise ValueError("label2d expects a 2D array")
        h, w = img.shape
        out = np.empty((h, w), dtype=np.uint32)
        n = ctypes.c_uint32(0)
        t = ctypes.c_float(0)
        self._check(self.lib.lbl_label2d(self._ctx, _alg_id(alg), img.ctypes.data, h, w, img.strides[0], conn,
                                         out.ctypes.data, ctypes.byref(n), ctypes.byref(t)))
        return out, int(n.value), float(t.value)

    def label3d(self, vol: np.ndarray, alg="bleq", conn: int = 0) -> Tuple[np.ndarray, int, float]:
        vol = np.ascontiguousarray(vol, dtype=np.uint8)
        if vol.ndim != 3:
            raise ValueError("label3d expects a 3D array")
        d, h, w = vol.shape
        out = np.empty((d, h, w), dtype=np.uint32)
        n = ctypes.c_uint32(0)
        t = ctypes.c_float(0)
        self._check(self.lib.lbl_label3d(self._ctx, _alg_id(alg), vol.ctypes.data, d, h, w, conn,
                                         out.ctypes.data, ctypes.byref(n), ctypes.byref(t)))
        return out, int(n.value), float(t.value)

    def label2d_dev(self, d_px: int, h: int, w: int, pitch: int, d_labels: int, alg="bleq", conn: int = 0,
                    d_ncomp: Optional[int] = None, stream: Optional[int] = None):
        self._check(self.lib.lbl_label2d_dev(self._ctx, _alg_id(alg), d_px, h, w, pitch, conn, d_labels,
                                             d_ncomp or None, stream or None))

    def label3d_dev(self, d_vox: int, d: int, h: int, w: int, d_labels: int, alg="bleq", conn: int = 0,
                    d_ncomp: Optional[int] = None, stream: Optional[int] = None):
        self._check(self.lib.lbl_label3d_dev(self._ctx, _alg_id(alg), d_vox, d, h, w, conn, d_labels,
                                             d_ncomp or None, stream or None))


def header_functions(path: str = HEADER_PATH):
    """Names of the functions declared in a public header (for export checks)."""
    import re

    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lbl_[a-z0-9_]+)\s*\(", text)))
