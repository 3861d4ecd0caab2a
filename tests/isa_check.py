"""Static checks of compiled gfx950 kernels (test infrastructure).

`undefined_reads(asm_text)` runs a must-defined dataflow analysis over each kernel's control-flow
graph in hipcc's assembly output and reports reads of storage that is not written on every path
from the kernel entry:

  * accumulation registers (AGPRs, a0..a255), which the register allocator uses as spill space in
    the kernels that need more than 256 vector registers;
  * scratch spill slots (scratch_load from an offset no scratch_store wrote);
  * lanes of the vector registers that hold spilled scalar registers (v_readlane with a constant
    lane that no v_writelane wrote).

A read of never-written storage is what a value that depends on whatever the GPU ran before
looks like in the machine code. Reads through a dynamic lane index (v_readlane with a scalar lane
operand, as in cross-lane reductions) are not spill traffic and are not tracked.

`wait_state_violations(lines)` checks the wait states that software (the compiler) must place
between dependent instructions on this target family, on every control-flow path:

  * a VALU instruction writes a scalar register (v_readlane / v_readfirstlane to an SGPR, the
    SGPR result of a compare or of a carry-out) and a vector memory instruction reads it as an
    address, resource or offset: 5 wait states;
  * the same write followed by v_readlane / v_writelane using it as the lane select: 4;
  * a scalar write of M0 followed by an LDS DMA (global_load_lds / buffer_load ... lds), which
    takes its LDS address from M0: 1.

Each instruction in between counts one wait state, s_nop N counts N + 1.
"""
import collections
import re

_BRANCH = ("s_cbranch", "s_branch", "s_endpgm", "s_setpc")


def kernels(asm_text):
    """{symbol: [lines]} for every function body in the assembly."""
    out = {}
    cur = None
    for line in asm_text.split("\n"):
        m = re.match(r"^(_Z\S+):", line)
        if m:
            cur = m.group(1)
            out[cur] = []
            continue
        if cur is None:
            continue
        if line.startswith(".Lfunc_end"):
            cur = None
            continue
        out[cur].append(line)
    return out


def _regs(tok):
    m = re.match(r"^([av])\[(\d+):(\d+)\]$", tok)
    if m:
        return [f"{m.group(1)}{i}" for i in range(int(m.group(2)), int(m.group(3)) + 1)]
    return [tok] if re.match(r"^[av]\d+$", tok) else []


def _effects(instr):
    toks = instr.replace(",", " ").split()
    op = toks[0]
    defs, uses = set(), set()
    if op.startswith("scratch_store") or op.startswith("scratch_load"):
        m = re.search(r"offset:(\d+)", instr)
        off = int(m.group(1)) if m else 0
        width = 4 if "dwordx4" in op else (3 if "dwordx3" in op else (2 if "dwordx2" in op else 1))
        slots = {f"S{off + 4 * k}" for k in range(width)}
        if op.startswith("scratch_store"):
            defs |= slots
        else:
            uses |= slots
    elif op.startswith("v_writelane"):
        if re.match(r"^\d+$", toks[3]):
            defs.add(f"L{toks[1]}:{toks[3]}")
    elif op.startswith("v_readlane"):
        if re.match(r"^\d+$", toks[3]):
            uses.add(f"L{toks[2]}:{toks[3]}")
    elif len(toks) > 1:
        is_store = op.startswith(("global_store", "ds_write", "buffer_store", "flat_store",
                                  "scratch_store"))
        first = [r for r in _regs(toks[1]) if r.startswith("a")]
        if is_store:
            uses.update(first)
        else:
            defs.update(first)
        for tok in toks[2:]:
            uses.update(r for r in _regs(tok) if r.startswith("a"))
    return defs, uses


def _blocks(lines):
    blocks = []
    cur = ("entry", [])
    anon = 0
    for line in lines:
        m = re.match(r"^(\.LBB\d+_\d+):", line)
        if m:
            blocks.append(cur)
            cur = (m.group(1), [])
            continue
        t = line.split(";")[0].strip()
        if not t or t.startswith("."):
            continue
        cur[1].append(t)
        if t.split()[0].startswith(_BRANCH):
            blocks.append(cur)
            anon += 1
            cur = (f"#{anon}", [])
    blocks.append(cur)
    return blocks


def undefined_reads(lines):
    """List of (block label, storage, instruction) for one kernel body."""
    blocks = _blocks(lines)
    index = {label: i for i, (label, _) in enumerate(blocks)}
    succ = collections.defaultdict(list)
    for i, (_, ins) in enumerate(blocks):
        last = ins[-1] if ins else ""
        op = last.split()[0] if last else ""
        if op.startswith("s_cbranch"):
            succ[i].append(index[last.split()[-1]])
            if i + 1 < len(blocks):
                succ[i].append(i + 1)
        elif op == "s_branch":
            succ[i].append(index[last.split()[-1]])
        elif op.startswith(("s_endpgm", "s_setpc")):
            pass
        elif i + 1 < len(blocks):
            succ[i].append(i + 1)
    pred = collections.defaultdict(list)
    for i, ss in succ.items():
        for j in ss:
            pred[j].append(i)
    eff = [[_effects(t) for t in ins] for _, ins in blocks]
    out = [None] * len(blocks)  # None: not yet reached (top of the must-lattice)

    def block_in(i):
        if i == 0:
            return set()
        ps = [out[p] for p in pred[i] if out[p] is not None]
        return set.intersection(*ps) if ps else None

    changed = True
    while changed:
        changed = False
        for i in range(len(blocks)):
            s = block_in(i)
            if s is None:
                continue
            for d, _ in eff[i]:
                s |= d
            if out[i] != s:
                out[i] = s
                changed = True
    found = []
    for i, (label, ins) in enumerate(blocks):
        s = block_in(i)
        if s is None:
            continue
        for (d, u), t in zip(eff[i], ins):
            for x in sorted(u - s):
                found.append((label, x, t))
            s |= d
    return found


def _sregs(tok):
    m = re.match(r"^s\[(\d+):(\d+)\]$", tok)
    if m:
        return {f"s{i}" for i in range(int(m.group(1)), int(m.group(2)) + 1)}
    if re.match(r"^s\d+$", tok):
        return {tok}
    if tok == "vcc":
        return {"vcc_lo", "vcc_hi"}
    if tok in ("vcc_lo", "vcc_hi", "m0"):
        return {tok}
    return set()


def _toks(t):
    return [x for x in t.replace(",", " ").split() if x]


def _valu_sgpr_defs(t):
    ts = _toks(t)
    op = ts[0]
    if not op.startswith("v_"):
        return set()
    if op.startswith(("v_readlane", "v_readfirstlane")):
        return _sregs(ts[1])
    if op.startswith(("v_cmp", "v_cmpx")):
        return _sregs(ts[1]) if op.endswith("_e64") else {"vcc_lo", "vcc_hi"}
    if op.startswith(("v_add_co", "v_sub_co", "v_subrev_co", "v_addc_co", "v_subb_co",
                      "v_subbrev_co", "v_mad_u64_u32", "v_mad_i64_i32", "v_div_scale")):
        d = _sregs(ts[2]) if len(ts) > 2 else set()
        if op.endswith("_e32"):
            d |= {"vcc_lo", "vcc_hi"}
        return d
    return set()


def _salu_m0_def(t):
    ts = _toks(t)
    return {"m0"} if ts[0].startswith("s_") and len(ts) > 1 and ts[1] == "m0" else set()


def _wait_states(t):
    ts = _toks(t)
    return int(ts[1], 0) + 1 if ts[0] == "s_nop" else 1


def _cfg_preds(blocks):
    index = {label: i for i, (label, _) in enumerate(blocks)}
    pred = collections.defaultdict(list)
    for i, (_, ins) in enumerate(blocks):
        last = ins[-1] if ins else ""
        op = last.split()[0] if last else ""
        if op.startswith("s_cbranch"):
            pred[index[last.split()[-1]]].append(i)
            if i + 1 < len(blocks):
                pred[i + 1].append(i)
        elif op == "s_branch":
            pred[index[last.split()[-1]]].append(i)
        elif op.startswith(("s_endpgm", "s_setpc")):
            pass
        elif i + 1 < len(blocks):
            pred[i + 1].append(i)
    return pred


def wait_state_violations(lines):
    """List of (rule, wait states found, producer, consumer, block label) for one kernel body."""
    blocks = _blocks(lines)
    pred = _cfg_preds(blocks)

    def back(bi, pos, regs, need, defs_of, acc, seen):
        # most recent producer of `regs` on every path back from instruction pos of block bi
        ins = blocks[bi][1]
        for j in range(pos - 1, -1, -1):
            if defs_of(ins[j]) & regs:
                return [(acc, ins[j])] if acc < need else []
            acc += _wait_states(ins[j])
            if acc >= need:
                return []
        out = []
        for p in pred[bi]:
            if (p, acc) not in seen:
                seen.add((p, acc))
                out += back(p, len(blocks[p][1]), regs, need, defs_of, acc, seen)
        return out

    found = []
    for bi, (label, ins) in enumerate(blocks):
        for k, t in enumerate(ins):
            ts = _toks(t)
            op = ts[0]
            checks = []
            if op.startswith(("global_", "buffer_", "scratch_", "flat_")):
                u = set()
                for x in ts[1:]:
                    u |= _sregs(x)
                u.discard("m0")
                if u:
                    checks.append(("VALU-SGPR->VMEM", u, 5, _valu_sgpr_defs))
            if op.startswith(("v_readlane", "v_writelane")) and len(ts) > 3 and _sregs(ts[3]):
                checks.append(("VALU-SGPR->lane-select", _sregs(ts[3]), 4, _valu_sgpr_defs))
            if op.startswith("global_load_lds") or (op.startswith("buffer_load") and
                                                     " lds" in t):
                checks.append(("SALU-M0->LDS-DMA", {"m0"}, 1, _salu_m0_def))
            for rule, regs, need, defs_of in checks:
                for ws, prod in back(bi, k, regs, need, defs_of, 0, set()):
                    found.append((rule, ws, prod, t, label))
    return found
