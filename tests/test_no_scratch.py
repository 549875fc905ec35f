"""Register-spill guard for the built library (CPU only): every gfx950 kernel
in libglsx.so runs without private scratch memory.  A kernel that spills
keeps its results but loses occupancy and memory bandwidth, so a build that
starts to spill is a regression.

The device code objects are the clang offload bundles in the library's
.hip_fatbin section; each kernel's metadata note records its
.private_segment_fixed_size (bytes of scratch per work-item)."""
import os
import shutil
import struct
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "glscv_amd", "libglsx.so")
LLVM = "/opt/rocm/llvm/bin"
MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"


def _tool(name):
    path = os.path.join(LLVM, name)
    if os.path.exists(path):
        return path
    found = shutil.which(name)
    if not found:
        pytest.skip(f"{name} not available")
    return found


def code_objects(fatbin: bytes, arch: str = "gfx950"):
    """The device code objects for `arch` in a concatenation of offload
    bundles (magic, entry count, then offset/size/triple per entry)."""
    out = []
    start = fatbin.find(MAGIC)
    while start >= 0:
        (n,) = struct.unpack_from("<Q", fatbin, start + 24)
        p = start + 32
        for _ in range(n):
            off, size, tlen = struct.unpack_from("<QQQ", fatbin, p)
            p += 24
            triple = fatbin[p:p + tlen].decode()
            p += tlen
            if triple.endswith(arch):
                out.append(fatbin[start + off:start + off + size])
        start = fatbin.find(MAGIC, start + 1)
    return out


def kernel_scratch(notes: str):
    """(kernel name, private segment bytes) pairs from llvm-readelf --notes:
    within a kernel's metadata map the arguments come first, then the
    kernel's .name, then its .private_segment_fixed_size."""
    res, name = [], None
    for line in notes.splitlines():
        s = line.strip()
        if s.startswith(".name:"):
            name = s.split(":", 1)[1].strip()
        elif s.startswith(".private_segment_fixed_size:"):
            res.append((name, int(s.split(":", 1)[1])))
    return res


def test_no_kernel_uses_scratch(tmp_path):
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} is not built")
    fatbin = tmp_path / "fatbin.bin"
    subprocess.run([_tool("llvm-objcopy"), f"--dump-section=.hip_fatbin={fatbin}", LIB, os.devnull],
                   check=True, capture_output=True)
    cos = code_objects(fatbin.read_bytes())
    assert cos, "no gfx950 code object in libglsx.so"
    kernels = []
    for k, co in enumerate(cos):
        path = tmp_path / f"co{k}.elf"
        path.write_bytes(co)
        notes = subprocess.run([_tool("llvm-readelf"), "--notes", str(path)], check=True, capture_output=True,
                               text=True).stdout
        kernels += kernel_scratch(notes)
    assert len(kernels) > 50, f"only {len(kernels)} kernels found in the metadata"
    spilling = [(n, b) for n, b in kernels if b]
    assert not spilling, "kernels with scratch: " + ", ".join(f"{n} ({b} B)" for n, b in spilling[:10])


def test_metadata_parser_reports_a_spill():
    """The parser pairs each kernel's name with its scratch size, past the
    argument names that precede it."""
    notes = """    .args:
      - .name:           a
        .offset:         0
    .name:           _ZN4glsx3fooEv
    .private_segment_fixed_size: 92
    .name:           _ZN4glsx3barEv
    .private_segment_fixed_size: 0"""
    assert kernel_scratch(notes) == [("_ZN4glsx3fooEv", 92), ("_ZN4glsx3barEv", 0)]
