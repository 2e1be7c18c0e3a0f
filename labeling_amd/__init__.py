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
# lbl_alg: every ALG_LIST entry of the reference (main.cpp:103-130)
BLEQ, LBEQ, BIN, RUNEQ, HERUN, GRBLOCK, OCV = range(7)
ALGS = {"bleq": BLEQ, "lbeq": LBEQ, "bin": BIN, "runeq": RUNEQ, "he-run": HERUN, "gr-block": GRBLOCK, "ocv": OCV}
# lbl_label_mode
LABELS_ROOT, LABELS_REFERENCE, LABELS_CANONICAL, LABELS_REFERENCE_UNCROPPED = 0, 1, 2, 3

_lib = None
_synth = None


BATCH_AUTOTUNE, BATCH_NO_GRAPH = 1, 2


class ImageDesc(ctypes.Structure):
    """lbl_image2d"""
    _fields_ = [("px", ctypes.c_void_p), ("labels", ctypes.c_void_p), ("h", ctypes.c_uint32),
                ("w", ctypes.c_uint32), ("pitch_bytes", ctypes.c_size_t)]


class ServerConfig(ctypes.Structure):
    """lbl_server_config"""
    _fields_ = [("streams", ctypes.c_uint32), ("hbm_budget_bytes", ctypes.c_uint64), ("max_batch", ctypes.c_uint32)]


class Request(ctypes.Structure):
    """lbl_request"""
    _fields_ = [("alg", ctypes.c_int), ("conn", ctypes.c_int), ("d", ctypes.c_uint32), ("h", ctypes.c_uint32),
                ("w", ctypes.c_uint32), ("px", ctypes.c_void_p), ("labels", ctypes.c_void_p)]


class Completion(ctypes.Structure):
    """lbl_completion"""
    _fields_ = [("ticket", ctypes.c_uint64), ("status", ctypes.c_int), ("n_components", ctypes.c_uint32),
                ("t_submit", ctypes.c_double), ("t_start", ctypes.c_double), ("t_done", ctypes.c_double)]


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
        "lbl_label2d_dev_noalloc": (i32, [P, i32, P, u32, u32, sz, i32, P, P, P]),
        "lbl_set_profiling": (i32, [P, i32]),
        "lbl_last_kernel_times": (i32, [P, pf, ctypes.POINTER(ctypes.c_char_p), i32]),
        "lbl_batch_create": (i32, [P, i32, ctypes.POINTER(ImageDesc), u32, i32, P, ctypes.c_uint,
                                   ctypes.POINTER(P)]),
        "lbl_batch_run": (i32, [P, P]),
        "lbl_batch_config": (i32, [P, pu32, pu32, pu32, ctypes.POINTER(i32)]),
        "lbl_batch_destroy": (None, [P]),
        "lbl_label2d_each": (i32, [P, i32, ctypes.POINTER(ImageDesc), u32, i32, P, P]),
        "lbl_autotune2d": (i32, [P, i32, P, u32, u32, sz, P, P]),
        "lbl_tuning_save": (i32, [P, ctypes.c_char_p]),
        "lbl_tuning_load": (i32, [P, ctypes.c_char_p]),
        "lbl_server_create": (i32, [P, ctypes.POINTER(ServerConfig), ctypes.POINTER(P)]),
        "lbl_server_submit": (i32, [P, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)]),
        "lbl_server_poll": (i32, [P, ctypes.POINTER(Completion), u32, pu32]),
        "lbl_server_drain": (i32, [P]),
        "lbl_server_inject_faults": (i32, [P, u32]),
        "lbl_server_now": (ctypes.c_double, [P]),
        "lbl_server_destroy": (None, [P]),
        "lbl_host_alloc": (P, [sz]),
        "lbl_host_free": (None, [P]),
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
            raise LabelingError(EINVAL, f"unknown algorithm {alg!r}")
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

    def autotune2d(self, d_px: int, h: int, w: int, pitch: int, d_labels: int, stream: Optional[int] = None,
                   alg="bleq"):
        self._check(self.lib.lbl_autotune2d(self._ctx, _alg_id(alg), d_px, h, w, pitch, d_labels, stream or None))

    def tuning(self) -> dict:
        """The context's tuning table ({"alg/kind/size class": [tile, flatten, us]})."""
        import json
        import tempfile

        with tempfile.NamedTemporaryFile("r", suffix=".json") as f:
            self._check(self.lib.lbl_tuning_save(self._ctx, f.name.encode()))
            return json.load(open(f.name))

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

    def label3d(self, vol: np.ndarray, alg="bleq", conn: int = 0, padded: bool = False) -> Tuple[np.ndarray, int, float]:
        """padded: the output has the reference's padded shape (mode LABELS_REFERENCE_UNCROPPED)."""
        vol = np.ascontiguousarray(vol, dtype=np.uint8)
        if vol.ndim != 3:
            raise ValueError("label3d expects a 3D array")
        d, h, w = vol.shape
        out = np.empty(tuple(ref_pad3(n) for n in (d, h, w)) if padded else (d, h, w), dtype=np.uint32)
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


class Batch:
    """lbl_batch: device images labelled together by one hipGraph launch."""

    def __init__(self, labeler: Labeler, images, alg="bleq", conn: int = 0, d_ncomp: Optional[int] = None,
                 flags: int = 0):
        """images: sequence of (d_px, d_labels, h, w, pitch_bytes) device descriptors."""
        self.labeler = labeler
        lib = labeler.lib
        n = len(images)
        descs = (ImageDesc * n)()
        for i, (px, lab, h, w, pitch) in enumerate(images):
            descs[i] = ImageDesc(px, lab, h, w, pitch)
        self._b = ctypes.c_void_p()
        labeler._check(lib.lbl_batch_create(labeler.handle, _alg_id(alg), descs, n, conn, d_ncomp or None, flags,
                                            ctypes.byref(self._b)))

    def run(self, stream: Optional[int] = None):
        self.labeler._check(self.labeler.lib.lbl_batch_run(self._b, stream or None))

    def config(self):
        t, tw, th, fl = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_int()
        self.labeler.lib.lbl_batch_config(self._b, ctypes.byref(t), ctypes.byref(tw), ctypes.byref(th),
                                          ctypes.byref(fl))
        return {"tile_id": t.value, "tile_blocks": [tw.value, th.value], "flatten": bool(fl.value)}

    def close(self):
        if self._b:
            self.labeler.lib.lbl_batch_destroy(self._b)
            self._b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PinnedArray:
    """numpy view of page-locked host memory (lbl_host_alloc)."""

    def __init__(self, shape, dtype):
        self.lib = load()
        self.shape, self.dtype = tuple(shape), np.dtype(dtype)
        nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = self.lib.lbl_host_alloc(max(nbytes, 1))
        if not self.ptr:
            raise MemoryError("lbl_host_alloc failed")
        buf = (ctypes.c_uint8 * max(nbytes, 1)).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype=np.uint8, count=nbytes).view(self.dtype).reshape(self.shape)

    def free(self):
        if self.ptr:
            self.array = None
            self.lib.lbl_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Server:
    """lbl_server: continuous-batching request server."""

    def __init__(self, labeler: Labeler, streams: int = 4, hbm_budget_bytes: int = 8 << 30, max_batch: int = 32):
        self.labeler = labeler
        self.lib = labeler.lib
        cfg = ServerConfig(streams, hbm_budget_bytes, max_batch)
        self._s = ctypes.c_void_p()
        labeler._check(self.lib.lbl_server_create(labeler.handle, ctypes.byref(cfg), ctypes.byref(self._s)))

    def submit(self, px_ptr: int, labels_ptr: int, shape, alg="bleq", conn: int = 0) -> int:
        if len(shape) == 2:
            d, (h, w) = 0, shape
        else:
            d, h, w = shape
        req = Request(_alg_id(alg), conn, d, h, w, px_ptr, labels_ptr)
        t = ctypes.c_uint64()
        st = self.lib.lbl_server_submit(self._s, ctypes.byref(req), ctypes.byref(t))
        if st != OK:
            raise LabelingError(st, "lbl_server_submit rejected the request")
        return t.value

    def poll(self, cap: int = 1024):
        out = (Completion * cap)()
        n = ctypes.c_uint32()
        st = self.lib.lbl_server_poll(self._s, out, cap, ctypes.byref(n))
        if st != OK:
            raise LabelingError(st, "server device failure")
        return [(c.ticket, c.status, c.n_components, c.t_submit, c.t_start, c.t_done) for c in out[:n.value]]

    def drain(self):
        st = self.lib.lbl_server_drain(self._s)
        if st != OK:
            raise LabelingError(st, "server device failure")

    def inject_faults(self, n: int):
        """Testing knob: the next n launched groups fail at their kernel launch."""
        self.labeler._check(self.lib.lbl_server_inject_faults(self._s, n))

    def now(self) -> float:
        return float(self.lib.lbl_server_now(self._s))

    def close(self):
        if self._s:
            self.lib.lbl_server_destroy(self._s)
            self._s = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ref_pad2(n: int) -> int:
    """The reference's 2D padded size ((n >> 5) << 5) + 32 (LabelingTools.cpp:103)."""
    return ((n >> 5) << 5) + 32


def ref_pad3(n: int) -> int:
    """The reference's 3D padded size (((n + 4) >> 5) << 5) + 32, data at +2 (LabelingTools.cpp:160-176)."""
    return (((n + 4) >> 5) << 5) + 32


def batched_workload(n: int, seed: int, lo: int = 64, hi: int = 2048):
    """Shapes and densities of the batched configuration: H, W ~ U{lo..hi} and
    p ~ U[0.05, 0.95] independently per image, from (seed, image index)."""
    span = hi - lo + 1
    out = []
    for i in range(n):
        h = lo + synth_u64(seed, 0xB, 3 * i) % span
        w = lo + synth_u64(seed, 0xB, 3 * i + 1) % span
        p = 0.05 + 0.9 * (synth_u64(seed, 0xB, 3 * i + 2) >> 11) / float(1 << 53)
        out.append((h, w, p))
    return out


def serving_request(i: int, seed: int):
    """Input i of the serving configuration's request stream: 80% 2D squares of
    side U{256..4096}, 20% 3D cubes of side U{64..256}; half of each are
    Gaussian-smoothed noise blobs (2D sigma 4, 50% quantile; 3D sigma 3, 40%
    quantile), half Bernoulli (2D p ~ U[0.05, 0.95], 3D p ~ U[0.05, 0.35]),
    all from (seed, i)."""
    u = synth_u64(seed, 0x5A, i)
    blob = (u >> 8) & 1
    if u % 10 < 8:
        side = 256 + synth_u64(seed, 0x5B, i) % (4096 - 256 + 1)
        if blob:
            return blobs2d(side, side, 4.0, 0.5, seed=i)
        return bernoulli((side, side), 0.05 + 0.9 * (u % 1000) / 1000, seed, stream=i)
    side = 64 + synth_u64(seed, 0x5B, i) % (256 - 64 + 1)
    if blob:
        return blobs3d(side, side, side, 3.0, 0.4, seed=i)
    return bernoulli((side, side, side), 0.05 + 0.3 * (u % 1000) / 1000, seed, stream=i)


def header_functions(path: str = HEADER_PATH):
    """Names of the functions declared in a public header (for export checks)."""
    import re

    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lbl_[a-z0-9_]+)\s*\(", text)))
