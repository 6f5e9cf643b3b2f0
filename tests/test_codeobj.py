"""Static checks on the gfx950 code objects inside libweno_mi355.so (CPU).

The library's device code sits in the .hip_fatbin section as a sequence of
clang offload bundles (one per translation unit).  The test unbundles every
gfx950 code object and reads each kernel's metadata (llvm-readelf --notes):

  * the producer/consumer sweep kernels (WENO5 and WENO7, every storage
    type and weight branch) hold their per-lane coefficients in registers for the whole launch,
    so they must not carry a scratch frame beyond a few spilled producer
    registers (<= 32 bytes per lane).  An out-of-line call that takes the
    coefficient struct's address makes the compiler keep the coefficients in
    a 240-byte scratch frame written at every launch (63 MB of extra HBM
    writes per 200^3 x 8 sweep, 5 % of its time);
  * the fused consumer-step kernels of the JS weights, the WENO5 adjoint's
    tile kernel, the reshape / order-1 copy kernels and the batched kernel
    carry no scratch.
"""
import os
import re
import struct
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "wenostandalone_amd", "libweno_mi355.so")
LLVM = "/opt/rocm/llvm/bin"
MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"


def _kernels(tmp_path):
    """kernel symbol -> (private_segment_fixed_size, vgpr_count)."""
    objcopy, readelf = os.path.join(LLVM, "llvm-objcopy"), os.path.join(LLVM, "llvm-readelf")
    if not (os.path.exists(LIB) and os.path.exists(objcopy) and os.path.exists(readelf)):
        pytest.skip("library or LLVM tools not present")
    fat = tmp_path / "fatbin.bin"
    subprocess.run([objcopy, f"--dump-section=.hip_fatbin={fat}", LIB, str(tmp_path / "lib.so")], check=True)
    data = fat.read_bytes()
    out = {}
    for bi, start in enumerate(m.start() for m in re.finditer(re.escape(MAGIC), data)):
        # bundle header: magic, entry count, then (offset, size, triple size, triple) per entry
        (n,) = struct.unpack_from("<Q", data, start + len(MAGIC))
        p = start + len(MAGIC) + 8
        for _ in range(n):
            off, size, tl = struct.unpack_from("<QQQ", data, p)
            p += 24
            triple = data[p:p + tl].decode()
            p += tl
            if "gfx950" not in triple:
                continue
            elf = tmp_path / f"co{bi}.elf"
            elf.write_bytes(data[start + off:start + off + size])
            notes = subprocess.run([readelf, "--notes", str(elf)], check=True, capture_output=True,
                                   text=True).stdout
            name, scratch = None, None
            for line in notes.splitlines():
                m = re.match(r"\s+\.name:\s+(\S+)", line)
                if m:
                    name = m.group(1)
                m = re.match(r"\s+\.private_segment_fixed_size:\s+(\d+)", line)
                if m:
                    scratch = int(m.group(1))
                m = re.match(r"\s+\.vgpr_count:\s+(\d+)", line)
                if m and name is not None and scratch is not None:
                    out[name] = (scratch, int(m.group(1)))
                    name, scratch = None, None
    return out


@pytest.fixture(scope="module")
def kernels(tmp_path_factory):
    k = _kernels(tmp_path_factory.mktemp("codeobj"))
    assert k, "no gfx950 kernels found in the library"
    return k


def test_sweep_kernels_hold_no_coefficient_frame(kernels):
    sweeps = {n: v for n, v in kernels.items()
              if "13sweep5_kernel" in n or "13sweep7_kernel" in n or "20sweep5_branch_kernel" in n}
    # JS: 2 orders x 3 storage types x 2 layouts x copy/core x limiter (48);
    # TENO / WENO-Z / mapped: 3 branches x 2 orders x 2 layouts x copy/core (24)
    assert len(sweeps) >= 72, len(sweeps)
    heavy = {n: s for n, (s, _) in sweeps.items() if s > 32}
    assert not heavy, heavy


def test_headline_sweep_kernels_have_no_scratch(kernels):
    # WENO5-JS FP64 core-only sweeps along the contiguous and a strided axis (the
    # bench workload's x launch, and its y and z launches)
    for name in ("_ZN2wm13sweep5_kernelIdLi0ELb0ELb0EEEvNS_8TileArgsE",
                 "_ZN2wm13sweep5_kernelIdLi1ELb0ELb0EEEvNS_8TileArgsE"):
        assert name in kernels, name
        assert kernels[name][0] == 0, (name, kernels[name])
        # 4 waves per SIMD: two 512-thread workgroups per CU
        assert kernels[name][1] <= 128, (name, kernels[name])


def test_js_consumer_step_kernels_have_no_scratch(kernels):
    js = {n: v for n, v in kernels.items() if "fused_flux_kernelILi5ELi1E" in n or "fused_flux_kernelILi7ELi1E" in n}
    assert js
    assert all(s == 0 for s, _ in js.values()), js


def test_adjoint_tile_kernel_has_no_scratch(kernels):
    # the WENO5 adjoint's lanes hold 27 coefficients for the whole launch: at
    # 128 VGPRs they spilled and were reloaded every pass
    k = {n: v for n, v in kernels.items() if "vjp5_tile_kernel" in n}
    assert len(k) == 1, k
    assert all(s == 0 for s, _ in k.values()), k


def test_permute_and_batched_kernels_have_no_scratch(kernels):
    # reshape / order-1 copy: pure data movement; batched: the JS cells keep
    # every coefficient in registers (186-198 VGPRs, one 512-thread
    # workgroup per CU)
    perm = {n: v for n, v in kernels.items() if "permute_" in n}
    # flat (3 storage types x 3 pair loads), rows (3), tiles (3 x paired or not)
    assert len(perm) >= 18, sorted(perm)
    assert all(s == 0 for s, _ in perm.values()), perm
    batched = {n: v for n, v in kernels.items() if "14batched_kernel" in n}
    assert len(batched) == 2, batched
    assert all(s == 0 and r <= 256 for s, r in batched.values()), batched
