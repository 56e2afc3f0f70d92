"""Control-flow check of gfx950 assembly (compiler -S output) for one kernel.

A workgroup barrier (s_barrier) must be reached by every wave of the
workgroup the same number of times.  Inside a *divergent* loop -- one whose
latch masks lanes off (s_andn2_b64 exec, ...) and branches back while any
lane remains (s_cbranch_execnz, or s_cbranch_execz around a long jump) --
lanes of one wave iterate different numbers of times while the wave keeps
executing the loop, so a barrier in that loop breaks the workgroup.  This is
how an earlier persistent kernel, whose scheduler ran under
`if (threadIdx.x == 0)`, hung: lanes 1-63 re-ran the barrier and the stale
item while lane 0 waited, masked off, to pop the next one.

Loops are taken from the compiler's own annotations ("This Loop Header",
"Parent Loop", "in Loop: Header=").  `divergent_barrier_loops(asm, kernel)`
returns (header, latch line) for every divergent loop of `kernel` that holds
an s_barrier (an empty list is the pass condition).
"""
import re

_BLOCK = re.compile(r"^(?:(\.LBB\d+_\d+):|; %bb\.(\d+):)")
_HDR = re.compile(r"This (?:Inner )?Loop Header: Depth=(\d+)")
_PARENT = re.compile(r"Parent Loop BB(\d+_\d+) Depth=(\d+)")
_INLOOP = re.compile(r"in Loop: Header=BB(\d+_\d+) Depth=(\d+)")
_LONG_TARGET = re.compile(r"\(\.LBB(\d+_\d+)-\.Lpost_getpc\d+\)")


def kernel_lines(asm_text, kernel):
    """The lines of `kernel`'s function body (mangled name contains `kernel`)."""
    lines = asm_text.splitlines()
    pat = re.compile(r"^(_Z\S*" + re.escape(kernel) + r"\S*):")
    start = next((i for i, l in enumerate(lines) if pat.match(l)), None)
    if start is None:
        raise ValueError(f"kernel {kernel} not found")
    out = []
    for l in lines[start + 1:]:
        if l.startswith(".Lfunc_end"):
            break
        out.append(l)
    return out


def _instr(l):
    return l.split(";", 1)[0].strip()


def _blocks(lines):
    """[(name or None, first line, last line, comment text)] in layout order."""
    starts = [i for i, l in enumerate(lines) if _BLOCK.match(l)]
    if not starts or starts[0] != 0:
        starts = [0] + starts
    out = []
    for k, s in enumerate(starts):
        e = starts[k + 1] - 1 if k + 1 < len(starts) else len(lines) - 1
        m = _BLOCK.match(lines[s])
        name = m.group(1)[4:] if m and m.group(1) else None  # "8_115" for .LBB8_115
        com = [lines[s].split(";", 1)[1] if ";" in lines[s] else ""]
        j = s + 1
        while j <= e and lines[j].lstrip().startswith(";"):
            com.append(lines[j])
            j += 1
        out.append((name, s, e, "\n".join(com)))
    return out


def divergent_barrier_loops(asm_text, kernel):
    lines = kernel_lines(asm_text, kernel)
    blocks = _blocks(lines)
    parent = {}     # loop header -> enclosing loop header (or None)
    innermost = []  # per block: innermost enclosing loop header (or None)
    for name, s, e, com in blocks:
        h = _HDR.search(com)
        if h and name:
            depth = int(h.group(1))
            par = [p for p, d in _PARENT.findall(com) if int(d) == depth - 1]
            parent[name] = par[0] if par else None
            innermost.append(name)
            continue
        m = _INLOOP.search(com)
        innermost.append(m.group(1) if m else None)

    def enclosing(b):
        out, h = [], innermost[b]
        while h is not None and h not in out:
            out.append(h)
            h = parent.get(h)
        return out

    divergent = {}  # header -> latch line
    for b, (name, s, e, com) in enumerate(blocks):
        encl = enclosing(b)
        if not encl:
            continue
        body = [(i, _instr(lines[i])) for i in range(s, e + 1)]
        masks = any(ins.startswith("s_andn2_b64 exec, exec") for _, ins in body)
        for k, (i, ins) in enumerate(body):
            target, div = None, False
            if ins.startswith(("s_cbranch_execnz", "s_cbranch_execz")):
                target, div = ins.split()[1][4:], True
            elif ins.startswith("s_branch"):
                target, div = ins.split()[1][4:], masks
            elif ins.startswith("s_setpc_b64"):
                for _, prev in reversed(body[:k]):
                    m = _LONG_TARGET.search(prev)
                    if m:
                        target = m.group(1)
                        break
                # a long jump entered past an "s_cbranch_execz skip" runs with
                # some lanes masked: the far form of s_cbranch_execnz
                pb = b - 1
                while pb >= 0:
                    last = [x for x in (_instr(lines[j]) for j in range(blocks[pb][1],
                                                                        blocks[pb][2] + 1)) if x]
                    if last:
                        div = last[-1].startswith("s_cbranch_execz") or masks
                        break
                    pb -= 1
            if target is not None and target in encl and div:
                divergent.setdefault(target, i)
    bad = []
    for b, (name, s, e, com) in enumerate(blocks):
        if not any(_instr(lines[i]) == "s_barrier" for i in range(s, e + 1)):
            continue
        for h in enclosing(b):
            if h in divergent:
                bad.append(("BB" + h, divergent[h]))
    return sorted(set(bad))
