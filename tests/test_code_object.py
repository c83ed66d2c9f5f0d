"""The built library's gfx950 code objects: every kernel runs without
compiler-inserted VGPR spills.

A VGPR spill is stored with the EXEC mask of the point where the allocator
places it.  A value that every lane holds, spilled inside a divergent
branch, comes back as stale scratch in the lanes that were inactive there.
The implicit FMM's leaf-group kernel, held to 64 VGPRs, once spilled the
FP64 copy of G that way and returned NaN or zero accelerations for random
groups.  This test reads the kernel descriptors' metadata
(.vgpr_spill_count) from the offload bundles in libnbg.so with
llvm-readelf, so a spill fails on the CPU before any GPU run.  SGPR spills
(to VGPR lanes, independent of EXEC) are allowed, and so is the ORDERED
walks' private per-depth array, which is scratch by design and not a spill.
"""

import os
import re
import struct
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "n-body_amd", "libnbg.so")
READELF = "/opt/rocm/llvm/bin/llvm-readelf"
MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"


def _code_objects(data):
    """The amdgcn entries of every clang offload bundle in the library."""
    out = []
    pos = 0
    while True:
        i = data.find(MAGIC, pos)
        if i < 0:
            return out
        (n,) = struct.unpack_from("<Q", data, i + len(MAGIC))
        off = i + len(MAGIC) + 8
        for _ in range(n):
            o, size, tlen = struct.unpack_from("<QQQ", data, off)
            off += 24
            triple = data[off:off + tlen].decode()
            off += tlen
            if "amdgcn" in triple:
                out.append((triple, data[i + o:i + o + size]))
        pos = i + 1


def _kernels(elf_bytes, tmp_path, k):
    path = tmp_path / f"co{k}.elf"
    path.write_bytes(elf_bytes)
    notes = subprocess.run([READELF, "--notes", str(path)], check=True, capture_output=True, text=True).stdout
    kernels = []
    for block in notes.split("  - .")[1:]:
        name = re.search(r"\.name:\s+(\S+)", "." + block)
        spill = re.search(r"\.vgpr_spill_count:\s+(\d+)", block)
        if name and spill:
            kernels.append((name.group(1), int(spill.group(1))))
    return kernels


def test_no_kernel_spills_vgprs(tmp_path):
    if not os.path.exists(LIB):
        pytest.skip("libnbg.so is not built")
    if not os.path.exists(READELF):
        pytest.skip("llvm-readelf is not installed")
    objs = _code_objects(open(LIB, "rb").read())
    assert objs and all("gfx950" in t for t, _ in objs), "the library must carry gfx950 code objects"
    kernels = []
    for k, (_, elf) in enumerate(objs):
        kernels += _kernels(elf, tmp_path, k)
    names = [n for n, _ in kernels]
    assert any("k_walk_lbvh_fast" in n for n in names) and any("k_fmm_groups" in n for n in names)
    spilled = [(n, c) for n, c in kernels if c]
    assert not spilled, f"kernels with VGPR spills: {spilled}"
