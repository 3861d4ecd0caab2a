"""CPU checks of the compiled FP64 solve and adjoint-local kernels (gfx950 assembly from hipcc,
same flags as the library build): no kernel reads an accumulation register, a scratch spill slot
or a scalar-spill lane that is not written first on every path from the kernel entry.

These kernels need about 512 registers per lane: the allocator spills into the accumulation
registers and scratch. With the delay line of the z-march left uninitialized, the compiled code
copied never-written register pairs through that spill space every plane step; starting it at
zero (like the z-queues) leaves every carried register defined. The checker itself is pinned on
small hand-written programs that do and do not read undefined storage.

The same kernels are checked for the wait states the compiler must insert between dependent
instructions (a scalar register written by a vector instruction and then used as a memory
address or lane select, M0 written before an LDS DMA); that checker is pinned the same way."""
import os
import shutil
import subprocess

import pytest

from astaroth_amd import build as B
from tests import isa_check as I

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_checker_flags_undefined_reads():
    prog = """\
_Z1kv:
\tv_accvgpr_write_b32 a1, v0
\ts_cbranch_scc1 .LBB0_2
\tv_accvgpr_write_b32 a2, v0
\tscratch_store_dword off, v3, off offset:16
\tv_writelane_b32 v9, s4, 5
.LBB0_2:
\tv_accvgpr_read_b32 v1, a1
\tv_accvgpr_read_b32 v2, a2
\tscratch_load_dword v4, off, off offset:16
\tv_readlane_b32 s4, v9, 5
\tv_readlane_b32 s5, v9, s6
\ts_endpgm
.Lfunc_end0:
"""
    body = I.kernels(prog)["_Z1kv"]
    found = sorted(x for _, x, _ in I.undefined_reads(body))
    # a1 is written on both paths; a2, the spill slot and the lane only on the fall-through one;
    # the dynamic-lane read is not spill traffic
    assert found == ["Lv9:5", "S16", "a2"]
    ok = prog.replace("\ts_cbranch_scc1 .LBB0_2\n", "")
    assert I.undefined_reads(I.kernels(ok)["_Z1kv"]) == []


def test_checker_loop_carried():
    """A value defined at the end of a loop body and read at its top is undefined on entry."""
    prog = """\
_Z1lv:
\tv_accvgpr_write_b32 a3, v0
.LBB0_1:
\tv_accvgpr_read_b32 v1, a3
\tv_accvgpr_read_b32 v2, a4
\tv_accvgpr_write_b32 a4, v1
\ts_cbranch_scc0 .LBB0_1
\ts_endpgm
.Lfunc_end0:
"""
    found = [x for _, x, _ in I.undefined_reads(I.kernels(prog)["_Z1lv"])]
    assert found == ["a4"]


def test_wait_state_checker():
    """Producer and consumer in one block, across a branch, and with enough wait states."""
    prog = """\
_Z1wv:
\tv_readlane_b32 s4, v9, 5
\tv_add_f32_e32 v1, v2, v3
\tglobal_load_dword v1, v2, s[4:5]
\ts_mov_b32 m0, s6
\tglobal_load_lds_dword v3, s[8:9]
\tv_cmp_gt_f32_e64 s[10:11], v1, v2
\ts_cbranch_scc1 .LBB0_2
\ts_nop 4
.LBB0_2:
\tv_readlane_b32 s12, v4, s10
\tv_readlane_b32 s7, v9, 6
\ts_nop 3
\tglobal_load_dword v1, v2, s[6:7]
\ts_endpgm
.Lfunc_end0:
"""
    found = sorted((r, ws) for r, ws, _, _, _ in I.wait_state_violations(I.kernels(prog)["_Z1wv"]))
    # s4 read 1 wait state after its VALU write; M0 read at once by the DMA; s10 read as a lane
    # select right after the compare on the taken path (the fall-through path waits 5); s7 read
    # 4 wait states after its write (5 needed)
    assert found == [("SALU-M0->LDS-DMA", 0), ("VALU-SGPR->VMEM", 1), ("VALU-SGPR->VMEM", 4),
                     ("VALU-SGPR->lane-select", 1)]
    ok = (prog.replace("\ts_nop 3\n", "\ts_nop 4\n")
          .replace("\tglobal_load_dword v1, v2, s[4:5]\n",
                   "\ts_nop 3\n\tglobal_load_dword v1, v2, s[4:5]\n")
          .replace("\tglobal_load_lds_dword", "\ts_nop 0\n\tglobal_load_lds_dword")
          .replace("\ts_cbranch_scc1 .LBB0_2\n", ""))
    assert I.wait_state_violations(I.kernels(ok)["_Z1wv"]) == []


@pytest.fixture(scope="module")
def fp64_tiled_asm(tmp_path_factory):
    if shutil.which(B._hipcc()) is None and not os.path.exists(B._hipcc()):
        pytest.skip("hipcc not available")
    src = os.path.join(REPO, "astaroth_amd", "csrc", "solve_tiled.hip")
    out = tmp_path_factory.mktemp("isa") / "solve_tiled_fp64.s"
    cmd = ([B._hipcc()] + B.HIP_FLAGS + ["-DAC_DOUBLE_PRECISION=1"] + B.PRECISION_FLAGS["fp64"] +
           B.SOURCE_FLAGS.get("solve_tiled.hip", []) +
           ["-I", os.path.join(REPO, "include"), "--cuda-device-only", "-S", src, "-o", str(out)])
    cmd = [c for c in cmd if c != "-fPIC"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    ks = I.kernels(out.read_text())
    names = [k for k in ks if "solve_tiled_kernel" in k or "adjoint_local_tiled_kernel" in k]
    assert len(names) >= 8
    return {k: ks[k] for k in names}


def test_fp64_tiled_kernels_read_no_undefined_storage(fp64_tiled_asm):
    bad = {}
    for k, body in fp64_tiled_asm.items():
        found = I.undefined_reads(body)
        if found:
            bad[k] = found[:5]
    assert not bad, bad


def test_fp64_tiled_kernels_keep_wait_states(fp64_tiled_asm):
    """The software-placed wait states (VALU-written SGPR before a memory instruction or a lane
    select, M0 before an LDS DMA) are present on every path of every FP64 tiled kernel."""
    bad = {}
    for k, body in fp64_tiled_asm.items():
        found = I.wait_state_violations(body)
        if found:
            bad[k] = found[:5]
    assert not bad, bad
