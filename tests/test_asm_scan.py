"""Static checks over the engine's gfx950 assembly (CPU only).

The compiler's hazard recognizer and its scalar-condition-code (SCC) liveness do
not look inside inline asm, so the engine's asm blocks must themselves avoid:

* clobbering a reserved register (hipcc reports it as -Winline-asm: the clobber
  is not honoured and the register's live value may be lost);
* leaving SCC written by an asm block where compiler code reads it next (a
  branch or select of the compiler's would then test the asm's SCC);
* a VMEM instruction in an asm block reading an address SGPR that a VALU op
  wrote within the previous 5 wait states (the load may use the old value);
* an LDS DMA in an asm block issued in the wait state right after an M0 write.

The test compiles the engine's device code to assembly with the build's flags
and scans every kernel.  A planted-mistake check runs each scan on a small
synthetic listing with the hazard present.
"""

import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "onedal_amd", "csrc", "km_engine.hip")
HIPCC = "/opt/rocm/bin/hipcc"

# instructions that read SCC
SCC_READERS = re.compile(r"^(s_cbranch_scc[01]|s_cselect_b(32|64)|s_addc_u32|s_subb_u32|s_cmovk_i32|"
                         r"s_cmov_b(32|64))\b")
# SALU instructions that do not write SCC (everything else in s_* arithmetic does)
SALU_NO_SCC = re.compile(r"^(s_mov_|s_movk_|s_cmov|s_cselect|s_getpc|s_setpc|s_swappc|s_waitcnt|s_nop|"
                         r"s_barrier|s_sleep|s_sethalt|s_setprio|s_sendmsg|s_trap|s_icache|s_incperflevel|"
                         r"s_decperflevel|s_ttracedata|s_endpgm|s_branch|s_cbranch|s_load|s_buffer|s_store|"
                         r"s_dcache|s_memtime|s_memrealtime|s_getreg|s_setreg|s_set_gpr|s_denorm|s_round|"
                         r"s_version|s_inst_prefetch|s_clause|s_delay|s_mul_i32|s_mul_hi|s_mulk|s_sext|s_brev|"
                         r"s_ff0|s_ff1|s_flbit|s_bitset|s_pack_|s_bitreplicate|s_quadmask)")


def _compile_asm():
    out = os.path.join(tempfile.mkdtemp(prefix="km_asm_"), "engine.s")
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall",
           "--cuda-device-only", "-S", "-I", os.path.join(ROOT, "include"), SRC, "-o", out]
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert proc.returncode == 0, proc.stdout[-4000:]
    with open(out) as f:
        return proc.stdout, f.read()


@pytest.fixture(scope="module")
def compiled():
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not present")
    return _compile_asm()


def _instructions(text):
    """(instruction line, inside-asm flag) for every instruction, kernels in order;
    labels are kept as ('<label>', False), function starts as ('<func>', False)."""
    out = []
    in_asm = False
    for raw in text.splitlines():
        line = raw.strip()
        if line.startswith(";;#ASMSTART"):
            in_asm = True
            continue
        if line.startswith(";;#ASMEND"):
            in_asm = False
            continue
        if not line or line.startswith(";") or line.startswith("."):
            if re.match(r"^\.LBB\S*:", line):
                out.append(("<label>", False))
            continue
        if re.match(r"^[A-Za-z_][\w.$]*:", line):
            out.append(("<func>", False))
            continue
        line = line.split(";")[0].strip()
        if line:
            out.append((line, in_asm))
    return out


def _op(line):
    return line.split()[0]


def _writes_scc(line):
    op = _op(line)
    if not op.startswith("s_"):
        return False
    if op.startswith("s_cmp") or op.startswith("s_bitcmp"):
        return True
    return SALU_NO_SCC.match(op) is None


def scan_scc(insts):
    """Positions where an SCC value written by an asm block is read by the next
    SCC reader (no SCC write in between, straight-line order)."""
    bad = []
    for i, (line, in_asm) in enumerate(insts):
        if not in_asm or not _writes_scc(line):
            continue
        for j in range(i + 1, len(insts)):
            l2, a2 = insts[j]
            if l2 == "<func>":
                break
            if l2 == "<label>":
                continue
            if SCC_READERS.match(_op(l2)):
                bad.append((i, line, l2))
                break
            if _writes_scc(l2):
                break
            if _op(l2) in ("s_branch", "s_endpgm", "s_setpc_b64"):
                break
    return bad


def _sregs(operand):
    """SGPR numbers named by an operand such as s5 or s[4:5]."""
    m = re.match(r"^s\[(\d+):(\d+)\]$", operand)
    if m:
        return set(range(int(m.group(1)), int(m.group(2)) + 1))
    m = re.match(r"^s(\d+)$", operand)
    if m:
        return {int(m.group(1))}
    return set()


def _operands(line):
    parts = line.split(None, 1)
    if len(parts) < 2:
        return []
    return [p.strip() for p in parts[1].split(",")]


def _wait_states(line):
    op = _op(line)
    if op == "s_nop":
        ops = _operands(line)
        return int(ops[0], 0) + 1 if ops else 1
    return 1


def scan_vmem_sgpr(insts):
    """Asm VMEM instructions whose SGPR operands were written by a VALU op less
    than 5 wait states before."""
    bad = []
    for i, (line, in_asm) in enumerate(insts):
        op = _op(line)
        if not in_asm or not (op.startswith("global_") or op.startswith("buffer_")):
            continue
        regs = set()
        for o in _operands(line)[1:]:
            regs |= _sregs(o)
        if not regs:
            continue
        ws = 0
        for j in range(i - 1, -1, -1):
            l2, _ = insts[j]
            if l2 in ("<label>", "<func>"):
                break
            if _op(l2).startswith("v_"):
                ops = _operands(l2)
                if ops and _sregs(ops[0]) & regs:
                    bad.append((i, line, l2))
                    break
            ws += _wait_states(l2)
            if ws >= 5:
                break
    return bad


def scan_m0_lds_dma(insts):
    """Asm LDS DMA instructions issued right after an M0 write."""
    bad = []
    for i, (line, in_asm) in enumerate(insts):
        op = _op(line)
        if not in_asm or not ("_lds_" in op or op.endswith("_lds")):
            continue
        j = i - 1
        while j >= 0 and insts[j][0] == "<label>":
            j -= 1
        if j >= 0:
            prev = insts[j][0]
            ops = _operands(prev)
            if ops and ops[0] == "m0" and _op(prev) != "s_nop":
                bad.append((i, line, prev))
    return bad


def test_no_inline_asm_warnings(compiled):
    log, _ = compiled
    warn = [l for l in log.splitlines() if "Winline-asm" in l or "inline asm" in l.lower()]
    assert not warn, "\n".join(warn[:20])


def test_asm_scc_not_consumed(compiled):
    _, text = compiled
    insts = _instructions(text)
    assert any(a for _, a in insts), "no inline asm found: the scan would check nothing"
    bad = scan_scc(insts)
    assert not bad, bad[:10]


def test_asm_vmem_sgpr_hazard(compiled):
    _, text = compiled
    insts = _instructions(text)
    n_dma = sum(1 for l, a in insts if a and "global_load_lds" in l)
    assert n_dma > 0, "no asm LDS DMA found: the scan would check nothing"
    assert not scan_vmem_sgpr(insts)
    assert not scan_m0_lds_dma(insts)


def test_scans_catch_planted_hazards():
    """Each scan flags a listing with its hazard and passes the fixed listing."""
    scc_bad = """
k:
\ts_cmp_ge_i32 s4, s5
;;#ASMSTART
\ts_or_b32 s6, s7, 3
;;#ASMEND
\tv_and_or_b32 v1, v2, v3, s6
\ts_cbranch_scc0 .LBB0_1
"""
    scc_ok = scc_bad.replace("\tv_and_or_b32 v1, v2, v3, s6\n", "\tv_and_or_b32 v1, v2, v3, s6\n\ts_cmp_ge_i32 s4, s5\n")
    assert scan_scc(_instructions(scc_bad))
    assert not scan_scc(_instructions(scc_ok))
    vm_bad = """
k:
\tv_readfirstlane_b32 s8, v1
\tv_readfirstlane_b32 s9, v2
;;#ASMSTART
\tglobal_load_lds_dwordx4 v3, s[8:9]
;;#ASMEND
"""
    vm_ok = """
k:
\tv_readfirstlane_b32 s8, v1
\tv_readfirstlane_b32 s9, v2
\ts_mov_b32 m0, s10
;;#ASMSTART
\ts_mov_b64 s[12:13], s[8:9]
\tglobal_load_lds_dwordx4 v3, s[12:13]
;;#ASMEND
"""
    assert scan_vmem_sgpr(_instructions(vm_bad))
    assert not scan_vmem_sgpr(_instructions(vm_ok))
    m0_bad = """
k:
\ts_mov_b32 m0, s10
;;#ASMSTART
\tglobal_load_lds_dwordx4 v3, s[12:13]
;;#ASMEND
"""
    assert scan_m0_lds_dma(_instructions(m0_bad))
    assert not scan_m0_lds_dma(_instructions(vm_ok))
