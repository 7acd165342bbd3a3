"""Private (scratch) segment sizes of the built library's kernels.

Reads the AMDGPU code-object metadata of rapt_amd/lib/librapt_hip.so (the
gfx950 objects in its .hip_fatbin section, one offload bundle per source
file), parsed here in Python: nothing runs on a GPU and no tool is started.

What it guards: when a device function that receives the kernel's by-value
arguments (or a pointer to a local array) is left out of line, the compiler
copies the whole argument struct into each lane's scratch and reads it from
there in the hot loop.  The Particle/GC/Adaptive kernels for gridded fields
did this with the adiabaticity tests (768-784 bytes per lane, the gridded
Particle batch 1.76x slower), and the bounce kernels with the field-line
integrals and the drift functor (520-670 bytes).  Register spills under the
two-waves budget stay below 640 bytes in every kernel; a 696-byte KArgs copy
would not.
"""
import os
import re
import struct

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "rapt_amd", "lib", "librapt_hip.so")
TARGET = b"hipv4-amdgcn-amd-amdhsa--gfx950"
MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"
NT_AMDGPU_METADATA = 32


def _sections(elf):
    """(name, type, offset, size) of every section of a little-endian ELF64."""
    assert elf[:4] == b"\x7fELF" and elf[4] == 2 and elf[5] == 1
    shoff, = struct.unpack_from("<Q", elf, 0x28)
    shentsize, shnum, shstrndx = struct.unpack_from("<HHH", elf, 0x3A)
    raw = [struct.unpack_from("<IIQQQQ", elf, shoff + i * shentsize) for i in range(shnum)]
    strtab = raw[shstrndx][4]
    out = []
    for name, typ, _flags, _addr, off, size in raw:
        end = elf.index(b"\0", strtab + name)
        out.append((elf[strtab + name:end].decode(), typ, off, size))
    return out


def _kernel_metadata(code_object):
    """The amdhsa.kernels list of an AMDGPU code object's metadata note."""
    import msgpack
    for _name, typ, off, size in _sections(code_object):
        if typ != 7:  # SHT_NOTE
            continue
        p, end = off, off + size
        while p + 12 <= end:
            namesz, descsz, ntype = struct.unpack_from("<III", code_object, p)
            name = code_object[p + 12:p + 12 + namesz].rstrip(b"\0")
            d0 = p + 12 + ((namesz + 3) & ~3)
            if name == b"AMDGPU" and ntype == NT_AMDGPU_METADATA:
                return msgpack.unpackb(code_object[d0:d0 + descsz], raw=False)["amdhsa.kernels"]
            p = d0 + ((descsz + 3) & ~3)
    raise AssertionError("no AMDGPU metadata note")


def _private_sizes(lib_path):
    """Kernel name -> private segment bytes per lane, over every gfx950 code
    object in the library's .hip_fatbin section (one offload bundle per
    source file: magic, entry count, then offset, size, triple per entry)."""
    data = open(lib_path, "rb").read()
    fat = [(off, size) for name, _t, off, size in _sections(data) if name == ".hip_fatbin"]
    assert fat, "no .hip_fatbin section"
    off, size = fat[0]
    fb = data[off:off + size]
    starts = [m.start() for m in re.finditer(re.escape(MAGIC), fb)]
    assert starts, "no offload bundle in .hip_fatbin"
    sizes = {}
    for s in starts:
        n, = struct.unpack_from("<Q", fb, s + len(MAGIC))
        q = s + len(MAGIC) + 8
        for _ in range(n):
            eoff, esize, tlen = struct.unpack_from("<QQQ", fb, q)
            triple = fb[q + 24:q + 24 + tlen]
            q += 24 + tlen
            if triple == TARGET:
                for k in _kernel_metadata(fb[s + eoff:s + eoff + esize]):
                    sizes[k[".name"]] = int(k[".private_segment_fixed_size"])
    return sizes


@pytest.fixture(scope="module")
def sizes():
    if not os.path.exists(LIB):
        pytest.skip("librapt_hip.so is not built")
    pytest.importorskip("msgpack")
    return _private_sizes(LIB)


def test_every_kernel_has_metadata(sizes):
    # the tracer instantiations of every field model and the bounce kernels
    assert any(k.startswith("_ZN4rapt7k_traceILi9E") for k in sizes)
    assert any("k_bounce_center" in k for k in sizes)
    assert any("k_trace_gcmix" in k for k in sizes)


def test_gridded_tracers_and_bounce_kernels_have_no_argument_copy(sizes):
    # (the user-written field's instantiations, model 10, hold the
    # interpreter's register file instead: see the next test)
    picked = {k: v for k, v in sizes.items()
              if (k.startswith("_ZN4rapt7k_traceILi9E") or "bounce" in k) and "ILi10E" not in k}
    assert len(picked) >= 10
    big = {k: v for k, v in picked.items() if v > 64}
    assert not big, big


def test_user_field_kernels_hold_only_the_interpreter(sizes):
    """Model 10 (a traced user field): the interpreter's 32 double registers
    (256 bytes per lane) and its call frame, no argument copy."""
    picked = {k: v for k, v in sizes.items() if "ILi10E" in k}
    assert len(picked) >= 5
    big = {k: v for k, v in picked.items() if v > 256 + 320}
    assert not big, big


def test_no_kernel_holds_a_copy_of_its_arguments(sizes):
    big = {k: v for k, v in sizes.items() if v >= 640}
    assert not big, big
