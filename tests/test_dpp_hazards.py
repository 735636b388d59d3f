"""The grouped operator kernels broadcast entry values inside 16-lane rows
with DPP row_newbcast written as inline assembly (v_add_u32_dpp,
v_fmac_f32_dpp), which the compiler's hazard recognizer cannot see into.
The ISA requires two wait states between a vector ALU write of a VGPR and a
DPP read of it, and five between a vector ALU write of EXEC and a DPP
instruction.  This test compiles the solver for gfx950 and walks the
control-flow graph of every kernel backwards from each DPP instruction
(through branches and loop back edges, counting s_nop wait states) to check
that no such write comes too close.  A planted violation shows the check can
fail."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def _vregs(op):
    op = op.strip()
    m = re.match(r"^v\[(\d+):(\d+)\]$", op)
    if m:
        return set(range(int(m.group(1)), int(m.group(2)) + 1))
    m = re.match(r"^v(\d+)$", op)
    return {int(m.group(1))} if m else set()


def _functions(asm):
    fn, body = None, []
    for ln in asm.split("\n"):
        m = re.match(r"^(_Z\S+):", ln)
        if m:
            fn, body = m.group(1), []
            continue
        if fn is None:
            continue
        t = ln.split(";")[0].strip()
        if t.startswith(".Lfunc_end"):
            yield fn, body
            fn = None
            continue
        if not t or (t.startswith(".") and not t.endswith(":")):
            continue
        body.append(t)


def dpp_hazards(asm):
    """(number of row_newbcast instructions, list of (dpp, writer) pairs)"""
    n_dpp, found = 0, []
    for _, body in _functions(asm):
        labels = {t[:-1]: i for i, t in enumerate(body) if t.endswith(":")}
        preds = {i: [] for i in range(len(body))}
        for i, t in enumerate(body):
            op = t.split()[0]
            if op.startswith("s_cbranch") or op == "s_branch":
                tgt = t.split()[1]
                if tgt in labels:
                    preds[labels[tgt]].append(i)
            if op not in ("s_branch", "s_endpgm", "s_setpc_b64") and i + 1 < len(body):
                preds[i + 1].append(i)

        def walk(i, need_v, need_x, regs, seen, src):
            for p in preds[i]:
                if (p, need_v, need_x) in seen:
                    continue
                seen.add((p, need_v, need_x))
                t = body[p]
                if t.endswith(":"):
                    walk(p, need_v, need_x, regs, seen, src)
                    continue
                op = t.split()[0]
                if op == "s_nop":
                    k = int(t.split()[1], 0) + 1
                    if need_v - k > 0 or need_x - k > 0:
                        walk(p, need_v - k, need_x - k, regs, seen, src)
                    continue
                dst = t[len(op):].split(",")[0]
                if need_v > 0 and op.startswith("v_") and not op.startswith("v_cmp") and _vregs(dst) & regs:
                    found.append((src, t))
                if need_x > 0 and op.startswith("v_cmpx"):
                    found.append((src, t))
                if need_v > 1 or need_x > 1:
                    walk(p, need_v - 1, need_x - 1, regs, seen, src)

        for i, t in enumerate(body):
            if "row_newbcast" not in t:
                continue
            n_dpp += 1
            op = t.split()[0]
            walk(i, 2, 5, _vregs(t[len(op):].split(",")[1]), set(), t)
    return n_dpp, found


@pytest.fixture(scope="module")
def solver_asm(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp("asm") / "elp_solver.s"
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "--cuda-device-only", "-S", "-o", str(out),
           os.path.join(ROOT, "elpaso-core_amd", "csrc", "elp_solver.hip")]
    subprocess.run(cmd, check=True, capture_output=True, timeout=600)
    return out.read_text()


def test_no_dpp_hazards_in_solver(solver_asm):
    n, found = dpp_hazards(solver_asm)
    assert n > 0, "no row_newbcast instructions: the grouped kernels are gone or renamed"
    assert not found, found[:5]


def test_planted_hazards_are_found(solver_asm):
    # a VALU write of the broadcast source right before one DPP add, and a
    # VALU write of EXEC right before one DPP FMA
    m = re.search(r"v_add_u32_dpp (v\d+), (v\d+), (v\d+) row_newbcast:1 ", solver_asm)
    assert m
    planted = solver_asm[:m.start()] + f"v_mov_b32_e32 {m.group(2)}, 0\n\t" + solver_asm[m.start():]
    m2 = re.search(r"v_fmac_f32_dpp v\d+, v\d+, v\d+ row_newbcast:3 ", planted)
    assert m2
    planted = planted[:m2.start()] + "v_cmpx_gt_u32_e32 vcc, 0, v0\n\t" + planted[m2.start():]
    _, found = dpp_hazards(planted)
    assert any(w.startswith("v_mov_b32") for _, w in found)
    assert any(w.startswith("v_cmpx") for _, w in found)
