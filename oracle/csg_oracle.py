"""CPU restatement of the reference's tape compiler (TEST INFRASTRUCTURE ONLY:
imported by tests/, never by the product library).

Input: an expression DAG as a node list (op, a, b, value) with the op codes
of csg::Operator (vulkan/include/csg/expression.h:14-22), node identity =
list index, inputs before users.  Output: the tape words [n, instrs], the
constant pool and the slot count of csg::Tape::Tape (vulkan/src/csg/
tape.cpp:14-21).  Each step follows the reference function it names:

  constant_fold      ConstantFold + ConstantFoldStep (simplifier.cpp:31-104)
  three_way_compare  ThreeWayCompare + OpOrder (simplifier.cpp:107-180)
  normal_form        NormalForm + NormalFormStep (simplifier.cpp:183-215);
                     std::sort of two operands with comparator `cmp <= 0`
                     swaps them iff cmp(second, first) <= 0 (libstdc++'s
                     insertion sort, /usr/include/c++/11/bits/stl_algo.h:
                     1819-1871)
  merge_duplicates   MergeDuplicates + AreDuplicate (simplifier.cpp:218-259):
                     a linear scan of the unique nodes for the first match
  compile_nodes      Simplify, TopoSort, BuildConstantPool,
                     ComputeLiveliness, BuildInstrs (tape.cpp:23-228)

Nodes are Python objects here (identity = object), rebuilt at every pass as
the reference's TopoMap rebuilds shared_ptr nodes.  Constant folding uses
float32 arithmetic, GLM's min/max and the oracle's sin/cos/exp (the
library's transcendentals, pinned against libm in test_oracle.py).

Beyond the reference (which asserts): a bare y/z/t root gets COPY 0 <- axis,
constant-pool indices past 255 use the B byte (wide format).  A NaN made by
folding reads its own pool entry (the reference reads entry 0).
"""
from __future__ import annotations

import sys

import numpy as np

X, Y, Z, T, CONST, SIN, COS, EXP, NEG, SQRT, ADD, SUB, MUL, DIV, MIN, MAX = range(16)
TAPE_OP = {CONST: 0, SIN: 1, COS: 2, ADD: 3, SUB: 4, MUL: 5, DIV: 6, MIN: 7, MAX: 8,
           EXP: 9, NEG: 10, SQRT: 11}
F32 = np.float32


class N:
    __slots__ = ("op", "inputs", "c")

    def __init__(self, op, inputs=(), c=0.0):
        self.op, self.inputs, self.c = op, list(inputs), F32(c)

    def is_input(self):
        return self.op > CONST

    def is_const(self, v=None):
        return self.op == CONST and (v is None or self.c == F32(v))


def from_nodes(nodes, root):
    objs = []
    for r in nodes:
        op = int(r["op"])
        ins = []
        if op > CONST:
            ins.append(objs[int(r["a"])])
            if op >= ADD:
                ins.append(objs[int(r["b"])])
        objs.append(N(op, ins, float(r["value"]) if op == CONST else 0.0))
    return objs[root]


def topo_iter(root):
    """TopoIter (expression.cpp:298-314): DFS post-order, inputs in order."""
    seen, out = set(), []

    def dfs(e):
        if id(e) in seen:
            return
        seen.add(id(e))
        if e.is_input():
            for c in e.inputs:
                dfs(c)
        out.append(e)

    old = sys.getrecursionlimit()
    sys.setrecursionlimit(max(old, 200000))
    try:
        dfs(root)
    finally:
        sys.setrecursionlimit(old)
    return out


def topo_map(root, f):
    res = {}
    for e in topo_iter(root):
        res[id(e)] = f(e, [res[id(c)] for c in e.inputs])
    return res[id(root)]


def apply_op(op, args, fns):
    """ApplyOperator (expression.cpp:113-146)."""
    a = F32(args[0])
    b = F32(args[1]) if len(args) > 1 else None
    with np.errstate(all="ignore"):
        if op == SIN:
            return F32(fns["sin"](float(a)))
        if op == COS:
            return F32(fns["cos"](float(a)))
        if op == EXP:
            return F32(fns["exp"](float(a)))
        if op == NEG:
            return -a
        if op == SQRT:
            return F32(np.sqrt(a))
        if op == ADD:
            return a + b
        if op == SUB:
            return a - b
        if op == MUL:
            return a * b
        if op == DIV:
            return a / b
        if op == MIN:
            return b if b < a else a
        if op == MAX:
            return b if a < b else a
    raise ValueError(op)


def neg(e):
    return N(NEG, [e])


def fold_step(e, fns):
    """ConstantFoldStep (simplifier.cpp:31-94)."""
    if not e.is_input():
        return e
    ins = e.inputs
    consts = [i.c for i in ins if i.is_const()]
    if len(consts) == len(ins):
        return N(CONST, c=apply_op(e.op, consts, fns))
    if e.op == ADD:
        if ins[0].is_const(0):
            return ins[1]
        if ins[1].is_const(0):
            return ins[0]
        if ins[0].op == NEG and ins[1].op == NEG:
            return neg(N(ADD, [ins[0], ins[1]]))
        if ins[0].op == NEG:
            return N(SUB, [ins[1], ins[0].inputs[0]])
        if ins[1].op == NEG:
            return N(SUB, [ins[0], ins[1].inputs[0]])
    elif e.op == MUL:
        if ins[0].is_const(0) or ins[1].is_const(0):
            return N(CONST, c=0.0)
        if ins[0].is_const(1):
            return ins[1]
        if ins[1].is_const(1):
            return ins[0]
        if ins[0].is_const(-1):
            return fold_step(neg(ins[1]), fns)
        if ins[1].is_const(-1):
            return fold_step(neg(ins[0]), fns)
        if ins[0].op == NEG or ins[1].op == NEG:
            return N(MUL, [fold_step(neg(ins[0]), fns), fold_step(neg(ins[1]), fns)])
    elif e.op == SUB:
        if ins[0].is_const(0):
            return fold_step(neg(ins[1]), fns)
        if ins[1].is_const(0):
            return ins[0]
        if ins[1].op == NEG:
            return fold_step(N(ADD, [ins[0], ins[1].inputs[0]]), fns)
    elif e.op == DIV:
        if ins[0].is_const(0):
            return N(CONST, c=0.0)
        if ins[1].is_const(1):
            return ins[0]
        if ins[1].is_const(-1):
            return fold_step(neg(ins[0]), fns)
        if ins[1].is_const():
            with np.errstate(all="ignore"):
                r = F32(1.0) / ins[1].c
            return fold_step(N(MUL, [ins[0], N(CONST, c=r)]), fns)
    elif e.op == NEG:
        if ins[0].op == NEG:
            return ins[0].inputs[0]
        if ins[0].op == SUB:
            return fold_step(N(SUB, [ins[0].inputs[1], ins[0].inputs[0]]), fns)
    return e


def constant_fold(root, fns):
    """ConstantFold (simplifier.cpp:97-104)."""
    return topo_map(root, lambda e, ins: e if not e.is_input()
                    else fold_step(N(e.op, ins), fns))


def three_way_compare(a, b):
    """ThreeWayCompare (simplifier.cpp:133-180); OpOrder is the op code."""
    if a.op != b.op:
        return -1 if a.op < b.op else 1
    if a.op == CONST:
        return -1 if a.c < b.c else (1 if a.c > b.c else 0)
    if len(a.inputs) != len(b.inputs):
        return -1 if len(a.inputs) < len(b.inputs) else 1
    for x, y in zip(a.inputs, b.inputs):
        c = three_way_compare(x, y)
        if c != 0:
            return c
    return 0


def normal_form(root):
    def step(e, ins):
        if not e.is_input():
            return e
        if e.op in (ADD, MUL, MIN, MAX):
            if three_way_compare(ins[1], ins[0]) <= 0:
                ins = [ins[1], ins[0]]
        return N(e.op, ins)
    old = sys.getrecursionlimit()
    sys.setrecursionlimit(max(old, 200000))
    try:
        return topo_map(root, step)
    finally:
        sys.setrecursionlimit(old)


def are_duplicate(a, b):
    """AreDuplicate (simplifier.cpp:218-233)."""
    if a.op != b.op:
        return False
    if a.op <= T:
        return True
    if a.op == CONST:
        return a.c == b.c
    if len(a.inputs) != len(b.inputs):
        return False
    return all(x is y for x, y in zip(a.inputs, b.inputs))


def merge_duplicates(root):
    """MergeDuplicates (simplifier.cpp:236-259); the scan of the unique list
    is kept per op (nodes of other ops never match)."""
    root = normal_form(root)
    unique = {}

    def step(e, ins):
        ne = N(CONST, c=e.c) if e.op == CONST else N(e.op, ins)
        for u in unique.setdefault(ne.op, []):
            if are_duplicate(ne, u):
                return u
        unique[ne.op].append(ne)
        return ne
    return topo_map(root, step)


def compile_nodes(nodes, root, fns):
    """csg::Tape::Tape (tape.cpp:14-228).  Returns (words, consts, slot_count)."""
    r = merge_duplicates(constant_fold(from_nodes(nodes, root), fns))
    exprs = topo_iter(r)
    idx = {id(e): i for i, e in enumerate(exprs)}
    pool, cidx, nan_idx = [], {}, {}
    for e in exprs:
        if e.op == CONST:
            if np.isnan(e.c):
                # NaN != NaN: BuildConstantPool never finds it, so every
                # occurrence is pushed (tape.cpp:44-55); the instruction reads
                # its own entry (the reference's lookup reads entry 0)
                nan_idx[id(e)] = len(pool)
                pool.append(e.c)
                continue
            key = float(e.c) if e.c != 0 else 0.0      # unordered_map<float>: -0 == 0
            if key not in cidx:
                cidx[key] = len(pool)
                pool.append(e.c)
    live = [-1] * len(exprs)
    for i, e in enumerate(exprs):
        if e.is_input():
            for c in e.inputs:
                live[idx[id(c)]] = max(i, live[idx[id(c)]])
    axes = [None] * 4
    for e in exprs:
        if e.op <= T:
            axes[e.op] = e
    slots = list(axes)

    def current(e):
        for s, v in enumerate(slots):
            if v is e:
                return s
        raise AssertionError("value not in a slot")

    def release(e):
        for s, v in enumerate(slots):
            if v is e:
                slots[s] = None
                return

    words = []
    for i, e in enumerate(exprs):
        if e.op <= T:
            continue
        a = b = 0
        if e.op == CONST:
            ci = nan_idx[id(e)] if np.isnan(e.c) else cidx[float(e.c) if e.c != 0 else 0.0]
            a, b = ci & 0xFF, ci >> 8
        else:
            a = current(e.inputs[0])
            b = current(e.inputs[1]) if len(e.inputs) > 1 else 0
            for c in e.inputs:
                if live[idx[id(c)]] == i:
                    release(c)
        out = next((s for s, v in enumerate(slots) if v is None), None)
        if out is None:
            slots.append(None)
            out = len(slots) - 1
        slots[out] = e
        words.append(TAPE_OP[e.op] | out << 8 | a << 16 | b << 24)
    if Y <= r.op <= T:
        words.append(12 | r.op << 16)
    assert not words or (words[-1] >> 8) & 0xFF == 0
    return (np.array([len(words)] + words, dtype=np.uint32), np.array(pool, dtype=np.float32),
            len(slots))
