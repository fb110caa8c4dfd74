"""CPU checks of the level-0 schedule of the pass-parallel root kernel
(frep_root_schedule returns the schedule frep_tape_upload builds).

Structure: with the slots' contents simulated as values (the result of
instruction i, or input x/y/z/t), every operand of every pass reads the value
the tape's own sequential slot semantics give that operand
(vulkan/src/csg/tape.cpp:274-302, vulkan/shaders/tape/interval.comp:121-160),
all reads of a pass happen before its writes, and the result operand holds
slot 0's final value.  Shortening: the kernel's reverse walk over the passes
(read and clear the output slot's live flag; if it was set, mark the
instruction kept and set its chosen operands' flags), fed the oracle's
min/max choices for a cell, keeps exactly the instructions the oracle's
shortening keeps (vulkan/shaders/tape/shorten.comp:164-223) and rewrites
MIN/MAX to COPY the same way."""
import numpy as np
import pytest

from frep import lib, shapes
from frep.tape import Builder, Tape, compile_tape

OP_CONST, OP_MIN, OP_MAX, OP_COPY = 0, 7, 8, 12
BINARY = {3, 4, 5, 6, 7, 8}
NOP = 15
KCONST = 0x8000


def trig_tape():
    b = Builder()
    x, y, z, t = b.x(), b.y(), b.z(), b.t()
    g = b.add(b.mul(b.sin(b.mul(b.const(6.0), x)), b.cos(b.mul(b.const(5.0), y))),
              b.div(b.exp(b.mul(b.const(-1.5), b.square(z))), b.add(b.const(1.5), t)))
    s = shapes.sphere_sdf(b, (0.1, -0.2, 0.05), 0.6)
    f = b.max(b.min(b.sub(g, b.const(0.3)), s), b.neg(b.add(s, b.const(0.4))))
    return compile_tape(b, f)


def slot0_unwritten():
    words = np.array([2, 0 | (4 << 8), 3 | (5 << 8) | (1 << 16) | (4 << 24)], dtype=np.uint32)
    return Tape(words=words, consts=np.array([0.25], dtype=np.float32), slot_count=6)


TAPES = {
    "csg24": lambda: shapes.random_csg(24, 2),
    "spheres8": lambda: shapes.spheres_union(8, 1),
    "trig": trig_tape,
    "slot0_unwritten": slot0_unwritten,
    "union60": lambda: shapes.spheres_union(60, 1010),
    "union400": lambda: shapes.spheres_union(400, 1035),
    "gyroid": shapes.gyroid_blend,
}


def sequential_operands(words):
    """Per instruction, the values its operands hold when the tape runs in
    order: ('in', s) for input slot s, ('op', j) for instruction j's result."""
    n = int(words[0])
    cur = {s: ("in", s) for s in range(4)}
    ops = []
    for i in range(n):
        w = int(words[1 + i])
        op, o, a, b = w & 0xFF, (w >> 8) & 0xFF, (w >> 16) & 0xFF, w >> 24
        if op == OP_CONST:
            ops.append((op, None, None))
        else:
            ops.append((op, cur[a], cur[b] if op in BINARY else cur[a]))
        cur[o] = ("op", i)
    return ops, cur[0]


def simulate(words, info, sched, lpc):
    """Runs the schedule on values instead of numbers; returns the passes as
    lists of (op, out slot, instruction, A, B) and the constant table."""
    n, P = int(words[0]), info["passes"]
    ops, result = sequential_operands(words)
    table = []
    for k in range(info["consts"]):
        e = int(sched[info["c0"] + k])
        i, pool = e & 0xFFFF, e >> 16
        w = int(words[1 + i])
        assert w & 0xFF == OP_CONST and pool == (w >> 16) & 0xFFFF
        table.append(i)
    assert P % 4 == 0
    ent = sched[info["e0"]:].reshape(P, lpc, 2)
    content = {s: ("in", s) for s in range(4)}

    def value(v):
        if v & KCONST:
            return ("op", table[v & (KCONST - 1)])
        assert v in content, "read of a slot never written"
        return content[v]

    passes, done = [], set()
    for p in range(P):
        rows, mask, writes = [], 0, {}
        for x, y in ent[p]:
            x, y = int(x), int(y)
            op, so, i = x & 0xF, (x >> 4) & 0xFFF, x >> 16
            if op == NOP:
                continue
            assert op == int(words[1 + i]) & 0xFF and i not in done
            mask |= 1 << op
            want_op, va, vb = ops[i]
            assert value(y & 0xFFFF) == va, (p, i, "A")
            if want_op in BINARY:
                assert value(y >> 16) == vb, (p, i, "B")
            assert so not in writes and 4 <= so < info["slots"]
            writes[so] = ("op", i)
            rows.append((op, so, i, y & 0xFFFF, y >> 16))
            done.add(i)
        assert int(sched[info["m0"] + p]) == mask
        assert len(rows) <= lpc
        content.update(writes)                       # after every read of the pass
        passes.append(rows)
    assert value(info["res"]) == result
    # every non-CONST instruction the result depends on was evaluated
    need, stack = set(), [result]
    while stack:
        v = stack.pop()
        if v[0] == "op" and v[1] not in need:
            need.add(v[1])
            op, va, vb = ops[v[1]]
            if op != OP_CONST:
                stack += [va, vb]
    assert {i for i in need if ops[i][0] != OP_CONST} == done
    return passes, table


def kept_tape(words, info, passes, table, choice_of):
    """The kernel's reverse walk and compaction for one cell."""
    n = int(words[0])
    flags = [0] * n
    for i, c in choice_of.items():
        flags[i] = c
    live = [0] * info["slots"]

    def mark(v):
        if v & KCONST:
            flags[table[v & (KCONST - 1)]] = 4
        else:
            live[v] = 1

    mark(info["res"])
    for rows in reversed(passes):
        lv = [live[so] for (_, so, _, _, _) in rows]
        for (_, so, _, _, _) in rows:
            live[so] = 0
        for (op, so, i, a, b), l in zip(rows, lv):
            if not l:
                continue
            f = flags[i]
            flags[i] = f | 4
            sel = 3 if op in (3, 4, 5, 6) else (f & 3 if op in (OP_MIN, OP_MAX) else 1)
            if sel & 1:
                mark(a)
            if sel & 2:
                mark(b)
    out = []
    for i in range(n):
        if not flags[i] & 4:
            continue
        w = int(words[1 + i])
        op, o, bslot = w & 0xFF, (w >> 8) & 0xFF, w >> 24
        if op in (OP_MIN, OP_MAX):
            if flags[i] & 3 == 1:
                w = (w & ~0xFF) | OP_COPY
            elif flags[i] & 3 == 2:
                w = OP_COPY | (o << 8) | (bslot << 16) | (bslot << 24)
        out.append(w)
    return out


@pytest.mark.parametrize("lpc", [64, 16])
@pytest.mark.parametrize("name", sorted(TAPES))
def test_root_schedule_structure_and_shortening(oracle, name, lpc):
    tape = TAPES[name]()
    words = np.asarray(tape.words, dtype=np.uint32)
    info, sched = lib.root_schedule(words, lpc)
    assert info["scheduled"] == 1 and sched is not None
    assert info["cells_per_block"] in (8, 16)
    assert info["table_bytes"] >= 10 * info["consts"]
    n = int(words[0])
    assert info["cell_bytes"] >= 8 * info["slots"] + ((n + 3) & ~3) + info["slots"]
    passes, table = simulate(words, info, sched, lpc)
    mm = [i for i in range(n) if int(words[1 + i]) & 0xFF in (OP_MIN, OP_MAX)]
    # ambiguous cells of a 16^3 root over [-1, 1]^3 (t = 0.5 for the trig tape)
    t = 0.5 if name == "trig" else 0.0
    dim = 16
    cs = np.float32(2.0) / np.float32(dim)
    checked = 0
    for idx in range(0, dim ** 3, 7):
        x, y, z = idx // (dim * dim), (idx // dim) % dim, idx % dim
        lo = np.array([np.float32(-1.0) + np.float32(c) * cs for c in (x, y, z)], dtype=np.float32)
        hi = (lo + cs).astype(np.float32)
        (l, h), ch = oracle.eval_interval(tape, lo, hi, t)
        if l > 0 or h < 0:
            continue
        ref, reused = oracle.shorten(words, ch)
        got = kept_tape(words, info, passes, table, {i: int(ch[k]) for k, i in enumerate(mm)})
        if reused:
            assert len(got) == n
        else:
            assert got == [int(v) for v in ref[1:]], idx
        checked += 1
        if checked >= 60:
            break
    assert checked > 0


def test_root_schedule_declines_unsuitable_tapes():
    # a slot read before any write
    bad = np.array([1, 3 | (4 << 8) | (5 << 16) | (0 << 24)], dtype=np.uint32)
    info, sched = lib.root_schedule(bad)
    assert info["scheduled"] == 0 and sched is None
    # a dependency chain (one instruction per pass)
    b = Builder()
    v = b.x()
    for _ in range(12):
        v = b.neg(v)
    info, _ = lib.root_schedule(np.asarray(compile_tape(b, v).words, dtype=np.uint32))
    assert info["scheduled"] == 0
    # malformed arguments
    with pytest.raises(lib.FrepError):
        lib.root_schedule(np.array([5, 0], dtype=np.uint32))
    with pytest.raises(lib.FrepError):
        lib.root_schedule(np.array([1, 13], dtype=np.uint32))


def test_root_schedule_lds_limit():
    words = np.asarray(shapes.spheres_union(400, 1035).words, dtype=np.uint32)
    big, _ = lib.root_schedule(words, 64, 160 * 1024)
    assert big["scheduled"] == 1 and big["cells_per_block"] == 8
    small, _ = lib.root_schedule(words, 64, 32 * 1024)
    assert small["scheduled"] == 0
