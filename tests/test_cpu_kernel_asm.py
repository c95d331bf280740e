"""Checks on the gfx950 assembly of the transport kernel (compiled here with hipcc, no GPU).

The cooperative-fetch variants issue their record loads from one inline-asm block that
writes M0 (the LDS destination of each load) and the exec mask. M0 is a register the
compiler reserves for its own LDS DMA loads, so the block is only sound while the compiler
emits no other M0 use in those variants; the exec mask must be restored within the block.
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "oct-mps_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"
# transport_kernel<MULTI, LOG, SKIP, SETS_LDS, COOP> with COOP = true: single-set and the
# serving (multi-set, sets in LDS or in global memory) variants, unlogged and logged.
COOP_KERNELS = ("ILb0ELb0ELb0ELb0ELb1EEEvNS_5KArgsE", "ILb0ELb1ELb0ELb0ELb1EEEvNS_5KArgsE",
                "ILb1ELb0ELb0ELb1ELb1EEEvNS_5KArgsE", "ILb1ELb1ELb0ELb1ELb1EEEvNS_5KArgsE",
                "ILb1ELb0ELb0ELb0ELb1EEEvNS_5KArgsE", "ILb1ELb1ELb0ELb0ELb1EEEvNS_5KArgsE")


@pytest.fixture(scope="module")
def asm(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp("isa") / "transport.s"
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", f"-I{ROOT}/include", f"-I{CSRC}",
           "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "--cuda-device-only", "-S", "-o", str(out), os.path.join(CSRC, "transport.hip")]
    subprocess.run(cmd, check=True, capture_output=True)
    return out.read_text()


def kernel_body(asm, mangled_suffix):
    name = "_ZN6octmps12_GLOBAL__N_116transport_kernel" + mangled_suffix
    start = asm.index(name + ":")
    end = asm.index("; Occupancy", start)
    return asm[start:end]


@pytest.mark.parametrize("suffix", COOP_KERNELS)
def test_m0_is_written_only_by_the_fetch_blocks(asm, suffix):
    body = kernel_body(asm, suffix)
    blocks = re.findall(r";;#ASMSTART\n(.*?);;#ASMEND", body, flags=re.S)
    fetch = [b for b in blocks if "buffer_load_dwordx4" in b]
    # one fetch block at the loop top and one at its end; the serving variant also fetches
    # the photons it restores from the previous slice
    assert len(fetch) == (3 if suffix.startswith("ILb1") else 2)
    outside = re.sub(r";;#ASMSTART\n.*?;;#ASMEND", "", body, flags=re.S)
    m0_outside = [l for l in outside.splitlines() if re.search(r"\bm0\b", l) and not l.lstrip().startswith(";")]
    assert not m0_outside, m0_outside
    for b in fetch:
        lines = [l.strip() for l in b.strip().splitlines()]
        # exec saved first and restored last; five loads, each after its M0 and exec writes
        assert lines[0].startswith("s_mov_b64") and lines[0].endswith("exec")
        assert lines[-1].startswith("s_mov_b64 exec,")
        assert sum(l.startswith("buffer_load_dwordx4") and l.endswith("lds") for l in lines) == 5
        assert sum(l.startswith(("s_mov_b32 m0", "s_add_u32 m0")) for l in lines) == 5


@pytest.mark.parametrize("suffix", (COOP_KERNELS[0], COOP_KERNELS[2], COOP_KERNELS[4]))
def test_cooperative_kernel_does_not_spill(asm, suffix):
    tail_start = asm.index("_ZN6octmps12_GLOBAL__N_116transport_kernel" + suffix + ":")
    tail = asm[tail_start:asm.index("; Occupancy", tail_start) + 200]
    assert re.search(r"; ScratchSize: 0\b", tail), "the unlogged cooperative kernel spills registers"
    # 6 waves per SIMD single-set; the serving variants are built for 5, the serving grid's
    # occupancy (the LDS of the variant with the set table allows 5 blocks per CU).
    assert re.search(r"; Occupancy: %d\b" % (6 if suffix == COOP_KERNELS[0] else 5), tail)
