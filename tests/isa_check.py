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
