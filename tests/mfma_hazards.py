"""MFMA result-hazard check on gfx950 device assembly built from this repository's sources.

On gfx9-family GPUs an MFMA writes its accumulator several cycles after issue; a VALU read of
that accumulator (v_accvgpr_read / v_accvgpr_mov) must be separated from the MFMA by wait states,
which the compiler inserts.  In one form of the panel GEMM kernel (runtime tile counts inside an
LDS-pipelined loop) the compiler placed such a read one instruction after a branch out of the
block holding the last MFMA, and the read returned the old value.  This check walks every
control-flow path backwards from each accumulator read to the nearest MFMA writing the register
(a v_accvgpr_write of it ends the path) and reports paths that cross a branch with fewer issue
slots (one per instruction, s_nop N = N + 1) than the minimum below for that MFMA: at or under the
smallest cross-branch distances the compiler produces for these kernels in their working forms
(FP64 19, FP32 10, BF16 8) and far above the failing one (1)."""
from __future__ import annotations

import collections
import re

MIN_SLOTS = {"v_mfma_f64_16x16x4_f64": 18, "v_mfma_f32_16x16x4_f32": 8, "v_mfma_f32_16x16x32_bf16": 6}
DEFAULT_MIN = 6
_BR = re.compile(r"^s_(c?branch\w*)\s+(\.LBB\w+)")


def _agprs(spec: str):
    spec = spec.rstrip(",")
    m = re.match(r"a\[(\d+):(\d+)\]$", spec)
    if m:
        return set(range(int(m.group(1)), int(m.group(2)) + 1))
    m = re.match(r"a(\d+)$", spec)
    return {int(m.group(1))} if m else set()


def functions(text: str):
    """(symbol, instruction lines with block labels) of every kernel in the assembly."""
    cur, body = None, []
    for ln in text.split("\n"):
        m = re.match(r"^(_Z\S+):", ln)
        if m:
            if cur:
                yield cur, body
            cur, body = m.group(1), []
            continue
        if cur is None:
            continue
        s = ln.split(";")[0].strip()
        if s.startswith(".Lfunc_end"):
            yield cur, body
            cur = None
            continue
        if not s or (s.startswith(".") and not re.match(r"^\.LBB\w+:", s)):
            continue
        body.append(s)
    if cur:
        yield cur, body


def cross_branch_reads(body):
    """(read, mfma opcode, distance) for every accumulator read reached from an MFMA writer through
    at least one branch or block boundary."""
    blocks, labels, cur = [], {}, []
    for s in body:
        m = re.match(r"^(\.LBB\w+):", s)
        if m:
            if cur:
                blocks.append(cur)
                cur = []
            labels[m.group(1)] = len(blocks)
            continue
        cur.append(s)
        if s.startswith(("s_branch", "s_cbranch", "s_endpgm", "s_setpc")):
            blocks.append(cur)
            cur = []
    if cur:
        blocks.append(cur)
    preds = collections.defaultdict(list)
    for i, b in enumerate(blocks):
        last = b[-1] if b else ""
        m = _BR.match(last)
        if m and m.group(2) in labels:
            preds[labels[m.group(2)]].append(i)
        if not last.startswith(("s_branch", "s_endpgm", "s_setpc")) and i + 1 < len(blocks):
            preds[i + 1].append(i)

    def walk(bi, idx, regs, dist, crossed, seen, out):
        b = blocks[bi]
        for j in range(idx - 1, -1, -1):
            s = b[j]
            op = s.split()[0]
            if op.startswith("v_mfma") and _agprs(s.split()[1]) & regs:
                if crossed:
                    out.append((op, dist))
                return
            if op == "v_accvgpr_write_b32" and _agprs(s.split()[1]) & regs:
                return
            dist += int(s.split()[1], 0) + 1 if op == "s_nop" else 1
            if dist > 40:
                return
        for p in preds[bi]:
            if (p, dist) not in seen:
                seen.add((p, dist))
                walk(p, len(blocks[p]), regs, dist, True, seen, out)

    found = []
    for bi, b in enumerate(blocks):
        for j, s in enumerate(b):
            if s.startswith(("v_accvgpr_read_b32", "v_accvgpr_mov_b32")):
                out = []
                walk(bi, j, _agprs(s.split()[2]), 0, False, set(), out)
                found.extend((s, op, d) for op, d in out)
    return found


def hazards(asm_text: str):
    """(kernel, read, mfma opcode, distance) below the minimum for that MFMA."""
    bad = []
    for name, body in functions(asm_text):
        for s, op, d in cross_branch_reads(body):
            if d < MIN_SLOTS.get(op, DEFAULT_MIN):
                bad.append((name, s, op, d))
    return bad
