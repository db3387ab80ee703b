"""Resource rules of the built device code (CPU only: reads the metadata of the gfx950 code
objects inside pspaqr_public_amd/libspaqr_gpu.so, as llvm-readelf prints it).

Every kernel keeps its working set in registers and LDS: no private segment (scratch), no
dynamic stack, no spilled VGPRs, LDS within the 160 KiB of a gfx950 CU.  (Spilled SGPRs go to
VGPR lanes, not to memory, and are allowed.)  The FP64
instantiation of the column-layout panel QR broke exactly this rule (its 128-row segments do
not fit the register file: 329 VGPRs spilled to 888 bytes per lane of scratch in
k_tpqrt_coop<double>, which no shipped kernel uses), and it is the one configuration that has
faulted the GPU, so a kernel that starts to spill fails here, before it runs anywhere.
"""
import os
import re
import shutil
import subprocess
import tempfile

import pytest

from pspaqr_public_amd import build

LLVM = "/opt/rocm/llvm/bin"
MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"


def _kernels():
    so = os.path.join(os.path.dirname(build.__file__), "libspaqr_gpu.so")
    if not os.path.exists(so):
        pytest.skip("library not built")
    if not os.path.exists(os.path.join(LLVM, "clang-offload-bundler")):
        pytest.skip("ROCm LLVM tools not present")
    d = tempfile.mkdtemp()
    try:
        fat = os.path.join(d, "fatbin")
        subprocess.run([os.path.join(LLVM, "llvm-objcopy"), "--dump-section=.hip_fatbin=" + fat, so,
                        os.path.join(d, "copy.so")], check=True, capture_output=True)
        data = open(fat, "rb").read()
        offs = [m.start() for m in re.finditer(re.escape(MAGIC), data)]
        out = {}
        for i, o in enumerate(offs):
            b = os.path.join(d, "b%d" % i)
            open(b, "wb").write(data[o: offs[i + 1] if i + 1 < len(offs) else len(data)])
            subprocess.run([os.path.join(LLVM, "clang-offload-bundler"), "--unbundle", "--type=o",
                            "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", "--input=" + b, "--output=" + b + ".co"],
                           check=True, capture_output=True)
            notes = subprocess.run([os.path.join(LLVM, "llvm-readelf"), "--notes", b + ".co"], check=True,
                                   capture_output=True, text=True).stdout
            # one metadata map per kernel, from its .name to the next kernel's
            for blk in re.split(r"\n\s+- \.", notes)[1:]:
                m = re.search(r"^\s*\.?name:\s+(\S+)", blk, re.M)
                if not m or ".private_segment_fixed_size" not in blk:
                    continue
                f = lambda key: re.search(r"\.%s:\s+(\S+)" % key, blk).group(1)
                out[m.group(1)] = {k: f(k) for k in ("private_segment_fixed_size", "uses_dynamic_stack",
                                                     "vgpr_spill_count", "sgpr_spill_count",
                                                     "group_segment_fixed_size")}
        return out
    finally:
        shutil.rmtree(d, ignore_errors=True)


def test_no_kernel_uses_scratch_or_spills():
    """The rule holds for every hand-written kernel.  rocPRIM's onesweep radix sort (the
    symbolic phase's sorts, instantiated from the library's headers) keeps a small per-lane
    array in the private segment by its own design (tens of bytes, no spilled registers): it is
    held to no spills, no dynamic stack and at most 256 bytes instead."""
    ks = _kernels()
    assert len(ks) > 100  # every translation unit's kernels were found
    lib = {k: v for k, v in ks.items() if k.startswith("_ZN7rocprim")}
    own = {k: v for k, v in ks.items() if k not in lib}
    bad = {k: v for k, v in own.items()
           if v["private_segment_fixed_size"] != "0" or v["uses_dynamic_stack"] != "false"
           or v["vgpr_spill_count"] != "0"}
    assert not bad, bad
    bad = {k: v for k, v in lib.items()
           if int(v["private_segment_fixed_size"]) > 256 or v["uses_dynamic_stack"] != "false"
           or v["vgpr_spill_count"] != "0"}
    assert not bad, bad
    assert max(int(v["group_segment_fixed_size"]) for v in ks.values()) <= 160 * 1024
