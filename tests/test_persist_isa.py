"""The persistent kernel's control flow, read from its gfx950 assembly.

The build writes the kernel's assembly (mustard_amd/_build/mu_persist-gfx950.s,
compiler -S output).  Every loop of k_persistent that holds a workgroup
barrier must be wave-uniform (tests/isa_check.py explains why).  The checker
itself is pinned on two small listings: the shape of the earlier kernel that
hung (the first barrier at the head of a divergent loop) and its uniform
form.
"""
import os
import subprocess

import pytest

from tests.isa_check import divergent_barrier_loops

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ASM = os.path.join(ROOT, "mustard_amd", "_build", "mu_persist-gfx950.s")

# The worker loop of the earlier kernel, reduced to its control flow: the
# scheduler under `if (tid == 0)` and, after the optimizer merged the
# end-of-item and next-pop tests across the latch, an inner loop headed by the
# first barrier whose latch masks lanes off and branches back.
HUNG = """\
_ZN2mu12k_persistentENS_8ExecViewE:
; %bb.0:
\ts_branch .LBB8_3
.LBB8_2:                                ;   in Loop: Header=BB8_3 Depth=1
\ts_andn2_b64 exec, exec, s[6:7]
\ts_cbranch_execnz .LBB8_3
\ts_endpgm
.LBB8_3:                                ; =>This Loop Header: Depth=1
                                        ;     Child Loop BB8_115 Depth 2
\tv_cmp_eq_u32_e32 vcc, 0, v7
\ts_and_saveexec_b64 s[0:1], vcc
\tglobal_atomic_add v1, v61, v1, s[6:7] sc0
\ts_or_b64 exec, exec, s[0:1]
\ts_branch .LBB8_115
.LBB8_114:                              ;   in Loop: Header=BB8_115 Depth=2
\ts_andn2_b64 exec, exec, s[6:7]
\ts_cbranch_execnz .LBB8_115
\ts_branch .LBB8_2
.LBB8_115:                              ;   Parent Loop BB8_3 Depth=1
                                        ; =>  This Loop Header: Depth=2
\ts_barrier
\tds_read_b32 v0, v0
\ts_branch .LBB8_114
.Lfunc_end8:
"""

UNIFORM = HUNG.replace("\ts_andn2_b64 exec, exec, s[6:7]\n\ts_cbranch_execnz .LBB8_115",
                       "\ts_cmp_lg_u32 s3, 0\n\ts_cbranch_scc1 .LBB8_115") \
              .replace("\ts_andn2_b64 exec, exec, s[6:7]\n\ts_cbranch_execnz .LBB8_3",
                       "\ts_cmp_lg_u32 s4, 0\n\ts_cbranch_scc1 .LBB8_3")


def test_checker_flags_the_hung_shape():
    assert {h for h, _ in divergent_barrier_loops(HUNG, "k_persistent")} == {"BB8_115", "BB8_3"}
    assert divergent_barrier_loops(UNIFORM, "k_persistent") == []


def test_persistent_kernel_loops_with_barriers_are_uniform():
    if not os.path.exists(ASM):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "mustard_amd"), ASM], check=True)
    text = open(ASM).read()
    assert "s_barrier" in text
    assert divergent_barrier_loops(text, "k_persistent") == []
