"""The HIP kernels' gfx950 assembly has no accumulator read that follows an MFMA writing it, through
a branch, closer than that MFMA's result latency (see mfma_hazards.py), and no kernel of the GEMM
sources uses per-thread scratch memory (a spill, or a local array indexed at run time: every
thread's copy costs HBM writes and reads).  Builds device assembly with the library's own compiler
flags; CPU only."""
import re
import os
import shutil
import subprocess

import pytest

import mfma_hazards as mh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "imps_cpp_tools_amd", "csrc")


_ASM = {}


def _device_asm(src, tmp_path):
    if src not in _ASM:
        _ASM[src] = _compile(src, tmp_path)
    return _ASM[src]


def _compile(src, tmp_path):
    from imps_cpp_tools_amd import build
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        pytest.skip("hipcc not available")
    flags = [f for f in build.HIPCC_FLAGS if f not in ("-shared", "-fPIC")]
    out = str(tmp_path / (os.path.basename(src) + ".s"))
    cmd = [exe] + flags + ["-I", os.path.join(ROOT, "include"), "-S", "--offload-device-only", src, "-o", out]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-2000:]
    with open(out) as f:
        return f.read()


def test_checker_flags_a_read_right_after_a_branch():
    """The check itself: an accumulator read one slot after a branch from its MFMA is reported, the
    same read after enough wait states is not, and a v_accvgpr_write in between clears it."""
    head = ["_Zk:", "v_mfma_f32_16x16x32_bf16 a[8:11], v[18:21], v[2:5], a[8:11]", "s_cbranch_vccnz .LBB0_2",
            "v_mfma_f32_16x16x32_bf16 a[0:3], v[26:29], v[6:9], a[0:3]", ".LBB0_2:"]
    bad = head + ["v_accvgpr_read_b32 v24, a11", "s_endpgm", ".Lfunc_end0:"]
    # both paths: the taken branch (1 slot) and the fall-through past the other MFMA (2 slots)
    assert sorted(h[3] for h in mh.hazards("\n".join(bad))) == [1, 2]
    ok = head + ["s_nop 7", "v_accvgpr_read_b32 v24, a11", "s_endpgm", ".Lfunc_end0:"]
    assert mh.hazards("\n".join(ok)) == []
    cleared = head + ["v_accvgpr_write_b32 a11, v3", "v_accvgpr_read_b32 v24, a11", "s_endpgm", ".Lfunc_end0:"]
    assert mh.hazards("\n".join(cleared)) == []


@pytest.mark.parametrize("src", ["panel_kernels.hip", "gemm_kernels.hip"])
def test_no_mfma_result_reads_across_branches(src, tmp_path):
    asm = _device_asm(os.path.join(CSRC, src), tmp_path)
    assert "v_mfma" in asm
    bad = mh.hazards(asm)
    assert not bad, bad[:5]


@pytest.mark.parametrize("src", ["panel_kernels.hip", "gemm_kernels.hip"])
def test_no_scratch_memory(src, tmp_path):
    asm = _device_asm(os.path.join(CSRC, src), tmp_path)
    sizes = re.findall(r"; ScratchSize: (\d+)", asm)
    assert sizes and all(int(x) == 0 for x in sizes), sizes
    assert "scratch_store" not in asm and "scratch_load" not in asm
