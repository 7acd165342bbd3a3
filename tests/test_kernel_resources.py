"""Private (scratch) segment sizes of the built library's kernels.

Reads the AMDGPU code-object metadata of rapt_amd/lib/librapt_hip.so (the
gfx950 objects in its .hip_fatbin section, one offload bundle per source
file) with the ROCm LLVM tools, without running anything on a GPU.

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
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "rapt_amd", "lib", "librapt_hip.so")
LLVM = "/opt/rocm/llvm/bin"
BUNDLER = os.path.join(LLVM, "clang-offload-bundler")
READELF = os.path.join(LLVM, "llvm-readelf")
TARGET = "hipv4-amdgcn-amd-amdhsa--gfx950"
MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"


def _private_sizes(tmp):
    fat = os.path.join(tmp, "fat.bin")
    subprocess.run(["objcopy", "--dump-section", ".hip_fatbin=" + fat, LIB], check=True)
    data = open(fat, "rb").read()
    starts = [m.start() for m in re.finditer(re.escape(MAGIC), data)]
    assert starts, "no offload bundle in .hip_fatbin"
    sizes = {}
    for i, s in enumerate(starts):
        e = starts[i + 1] if i + 1 < len(starts) else len(data)
        b = os.path.join(tmp, "b%d.bin" % i)
        o = os.path.join(tmp, "d%d.o" % i)
        with open(b, "wb") as f:
            f.write(data[s:e])
        subprocess.run([BUNDLER, "--type=o", "--targets=" + TARGET, "--input=" + b,
                        "--output=" + o, "--unbundle"], check=True)
        notes = subprocess.run([READELF, "--notes", o], capture_output=True, text=True,
                               check=True).stdout
        name = None
        for line in notes.splitlines():
            m = re.match(r"\s+\.name:\s+(\S+)", line)
            if m:
                name = m.group(1)
            m = re.match(r"\s+\.private_segment_fixed_size:\s+(\d+)", line)
            if m and name:
                sizes[name] = int(m.group(1))
                name = None
    return sizes


@pytest.fixture(scope="module")
def sizes(tmp_path_factory):
    if not os.path.exists(LIB):
        pytest.skip("librapt_hip.so is not built")
    for tool in (BUNDLER, READELF):
        if not os.path.exists(tool):
            pytest.skip(tool + " is not installed")
    if shutil.which("objcopy") is None:
        pytest.skip("objcopy is not installed")
    return _private_sizes(str(tmp_path_factory.mktemp("kmeta")))


def test_every_kernel_has_metadata(sizes):
    # the tracer instantiations of every field model and the bounce kernels
    assert any(k.startswith("_ZN4rapt7k_traceILi9E") for k in sizes)
    assert any("k_bounce_center" in k for k in sizes)
    assert any("k_trace_gcmix" in k for k in sizes)


def test_gridded_tracers_and_bounce_kernels_have_no_argument_copy(sizes):
    picked = {k: v for k, v in sizes.items()
              if k.startswith("_ZN4rapt7k_traceILi9E") or "bounce" in k}
    assert len(picked) >= 10
    big = {k: v for k, v in picked.items() if v > 64}
    assert not big, big


def test_no_kernel_holds_a_copy_of_its_arguments(sizes):
    big = {k: v for k, v in sizes.items() if v >= 640}
    assert not big, big
