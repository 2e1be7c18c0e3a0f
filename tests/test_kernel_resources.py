"""Resource use of the built gfx950 kernels, read from the library's own code
object metadata (no GPU needed).

* No kernel uses scratch (private) memory.  A runtime index into a register
  array (a window of words chosen by a row's alignment, say) makes the
  compiler keep that array in scratch, and every access becomes a memory
  round trip; the 8192^2 final kernel went from 50 to 87 us when that
  happened.
* Kernels that declare no shared arrays of their own hold at most the 64
  bytes of the count partials in LDS: the compiler may also move such an
  array into LDS instead of scratch (one slot per thread), which costs LDS
  round trips on the store path.  The 2D mark kernel declares its two
  16-word key-bit windows (256 bytes) and nothing more.
* The single-image tile kernels fit 64 VGPRs (8 waves per SIMD, which their
  launch bounds ask for).
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "labeling_amd", "liblabeling_amd.so")
LLVM = "/opt/rocm/llvm/bin"


def _kernels(tmp_path):
    objdump, readelf = os.path.join(LLVM, "llvm-objdump"), os.path.join(LLVM, "llvm-readelf")
    if not (os.path.exists(LIB) and os.path.exists(objdump) and os.path.exists(readelf)):
        pytest.skip("library or LLVM tools not present")
    lib = tmp_path / "lib.so"
    shutil.copy(LIB, lib)
    # writes lib.so.<i>.<target> files beside its input
    subprocess.run([objdump, "--offloading", str(lib)], check=True, capture_output=True, cwd=tmp_path)
    objs = sorted(p for p in tmp_path.iterdir() if "amdgcn" in p.name and p.stat().st_size)
    assert objs, "no gfx950 code objects in the library"
    assert all("gfx950" in p.name for p in objs), [p.name for p in objs]
    kernels = {}
    for p in objs:
        notes = subprocess.run([readelf, "--notes", str(p)], check=True, capture_output=True, text=True).stdout
        # one "  - .field:" line opens each kernel's entry; its own fields sit
        # four spaces in, before and after its .name
        entries, cur = [], None
        for line in notes.splitlines():
            m = re.match(r"^  - \.(\w+):\s+(\S+)", line)
            if m:
                cur = {m.group(1): m.group(2)}
                entries.append(cur)
                continue
            m = re.match(r"^    \.(\w+):\s+(\S+)", line)
            if m and cur is not None:
                cur[m.group(1)] = m.group(2)
        for e in entries:
            if "name" in e:
                kernels[e["name"]] = {k: int(v) for k, v in e.items() if v.isdigit()}
    return kernels


def test_no_kernel_uses_scratch(tmp_path):
    kernels = _kernels(tmp_path)
    assert len(kernels) > 50, len(kernels)
    spills = {k: v["private_segment_fixed_size"] for k, v in kernels.items() if v.get("private_segment_fixed_size", 0)}
    assert not spills, spills


def test_single_image_tile_kernels_fit_64_vgprs(tmp_path):
    kernels = _kernels(tmp_path)
    tiles = {k: v["vgpr_count"] for k, v in kernels.items() if "bleq2d_tile" in k and "3One" in k}
    assert len(tiles) >= 7, tiles
    assert all(n <= 64 for n in tiles.values()), tiles


def test_no_register_arrays_in_lds(tmp_path):
    kernels = _kernels(tmp_path)
    plain = ("bleq2d_final", "bleq2d_keys", "bleq2d_mark", "bleq2d_flatten", "bleq3d_final", "bleq3d_keys",
             "bleq3d_mark", "lbeq_final", "lbeq_relabel", "lbeq2d_flatten")
    seen = {k: v.get("group_segment_fixed_size", 0) for k, v in kernels.items() if any(p in k for p in plain)}
    assert len(seen) >= 30, sorted(seen)
    declared = {"bleq2d_mark": 256}   # bytes of a kernel's own shared arrays
    over = {k: n for k, n in seen.items() if n > max([64] + [v for p, v in declared.items() if p in k])}
    assert not over, over
