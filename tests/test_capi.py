"""Host-side checks of the C ABI library (no kernel launches: these run without a GPU)."""
import ctypes
import hashlib
import json
import os

import numpy as np
import pytest

import labeling_amd as L

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_library_exports_every_header_symbol():
    lib = ctypes.CDLL(L.LIB_PATH)
    names = L.header_functions()
    assert "lbl_label2d" in names and "lbl_label3d_dev" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared in include/labeling_amd.h but not exported: {missing}"


def test_registry_names_mirror_alg_list():
    lib = L.load()
    out = ctypes.c_int(-1)
    # every ALG_LIST name of the reference (main.cpp:103-130)
    for name, alg in (("bleq", L.BLEQ), ("lbeq", L.LBEQ), ("bin", L.BIN), ("runeq", L.RUNEQ), ("he-run", L.HERUN),
                      ("gr-block", L.GRBLOCK), ("ocv", L.OCV)):
        out.value = -1
        assert lib.lbl_alg_from_name(name.encode(), ctypes.byref(out)) == L.OK and out.value == alg
        assert L.ALGS[name] == alg
    assert lib.lbl_alg_from_name(b"nope", ctypes.byref(out)) == L.EINVAL


def test_status_strings():
    lib = L.load()
    for i, name in enumerate(L.STATUS_NAMES):
        assert lib.lbl_status_string(i).decode() == name


def test_create_rejects_bad_arguments():
    lib = L.load()
    assert lib.lbl_create(None, 0) == L.EINVAL
    h = ctypes.c_void_p()
    st = lib.lbl_create(ctypes.byref(h), 1 << 20)
    assert st in (L.EINVAL, L.EDEVICE)
    assert lib.lbl_last_error(None)
    assert not h.value


def test_null_context_is_rejected():
    lib = L.load()
    assert lib.lbl_label2d(None, 0, None, 1, 1, 1, 0, None, None, None) == L.EINVAL
    assert lib.lbl_set_label_mode(None, 0) == L.EINVAL


def test_synth_is_deterministic_and_matches_golden_hash():
    a = L.bernoulli((1024, 1024), 0.5, 0)
    b = L.bernoulli((1024, 1024), 0.5, 0)
    np.testing.assert_array_equal(a, b)
    assert set(np.unique(a)) <= {0, 255}
    assert abs(np.count_nonzero(a) / a.size - 0.5) < 0.01
    path = os.path.join(GOLD, "large.json")
    if os.path.exists(path):
        entry = json.load(open(path))["c1_bern1024_p5_s0"]
        assert hashlib.sha256(a.tobytes()).hexdigest() == entry["input_sha256"]


def test_synth_blobs_quantile():
    img = L.blobs2d(128, 96, 4.0, 0.5, 1)
    frac = np.count_nonzero(img) / img.size
    assert abs(frac - 0.5) < 1e-3
    vol = L.blobs3d(20, 24, 28, 3.0, 0.4, 2)
    assert abs(np.count_nonzero(vol) / vol.size - 0.6) < 1e-3


def test_header_is_plain_c():
    """The boundary header compiles as C99 with no HIP/torch headers."""
    import subprocess
    import tempfile

    src = '#include "labeling_amd.h"\nint main(void){ lbl_ctx *c = 0; (void)c; return 0; }\n'
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "t.c")
        open(p, "w").write(src)
        inc = os.path.join(os.path.dirname(L.HEADER_PATH))
        r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", inc, "-c", p, "-o", os.path.join(d, "t.o")],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
