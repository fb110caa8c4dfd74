"""Expression DAG -> SSA tape, in the reference's tape format.

Tape words (u32): ``[n, instr_0, ..., instr_{n-1}]`` with
``instr = op | out << 8 | A << 16 | B << 24`` (vulkan/include/csg/tape.h:15-42,
vulkan/src/engine/voxelizer.cpp:385-427).  Slots 0..3 hold x, y, z, t; the
result is slot 0.  A CONST instruction reads ``pool[A | B << 8]``: the
reference's own tapes always have B = 0 for CONST (vulkan/src/csg/tape.cpp:172
zero-initialises the instruction), so reference tapes decode unchanged, and
the 16-bit index lets pools grow past the reference's 256 constants.

The compiler follows vulkan/src/csg/tape.cpp:14-228:
  * topological order = DFS post-order from the root, inputs left to right
    (vulkan/src/csg/expression.cpp:298-314);
  * the constant pool is filled in order of first appearance (tape.cpp:44-55);
  * liveness = index of the last instruction reading a value (tape.cpp:66-79);
  * slots 0..3 start holding x, y, z, t (empty if the axis is unused); an
    instruction first releases the slots of inputs whose last use it is, then
    takes the lowest free slot for its output (tape.cpp:151-228).
Structural sharing of identical sub-expressions (hash-consing at build time)
and folding of all-constant operations play the role of the reference's
MergeDuplicates and ConstantFold (vulkan/src/csg/simplifier.cpp:97-104,236-259).
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np

OP_CONST, OP_SIN, OP_COS, OP_ADD, OP_SUB, OP_MUL, OP_DIV = 0, 1, 2, 3, 4, 5, 6
OP_MIN, OP_MAX, OP_EXP, OP_NEG, OP_SQRT, OP_COPY = 7, 8, 9, 10, 11, 12
OP_NAMES = ["CONST", "SIN", "COS", "ADD", "SUB", "MUL", "DIV", "MIN", "MAX",
            "EXP", "NEG", "SQRT", "COPY"]

_AXES = ("x", "y", "z", "t")
_UNARY = {"sin": OP_SIN, "cos": OP_COS, "exp": OP_EXP, "neg": OP_NEG, "sqrt": OP_SQRT}
_BINARY = {"add": OP_ADD, "sub": OP_SUB, "mul": OP_MUL, "div": OP_DIV,
           "min": OP_MIN, "max": OP_MAX}


def f32(v) -> float:
    """Round a Python float to binary32 and back."""
    return float(np.float32(v))


def _apply(op: str, args):
    """Float32 semantics of one operation (used for constant folding)."""
    a = np.float32(args[0])
    b = np.float32(args[1]) if len(args) > 1 else None
    with np.errstate(all="ignore"):
        if op == "add":
            return float(a + b)
        if op == "sub":
            return float(a - b)
        if op == "mul":
            return float(a * b)
        if op == "div":
            return float(a / b)
        if op == "min":
            return float(b if b < a else a)
        if op == "max":
            return float(b if a < b else a)
        if op == "neg":
            return float(-a)
        if op == "sqrt":
            return float(np.sqrt(a))
    raise ValueError(f"cannot fold {op}")


class Builder:
    """Hash-consing expression builder.  Node ids are indices into ``nodes``."""

    def __init__(self):
        self.nodes: list[tuple] = []
        self._index: dict[tuple, int] = {}

    def _intern(self, key: tuple) -> int:
        i = self._index.get(key)
        if i is None:
            i = len(self.nodes)
            self.nodes.append(key)
            self._index[key] = i
        return i

    # leaves
    def axis(self, name: str) -> int:
        return self._intern(("axis", name))

    def x(self):
        return self.axis("x")

    def y(self):
        return self.axis("y")

    def z(self):
        return self.axis("z")

    def t(self):
        return self.axis("t")

    def const(self, v: float) -> int:
        v = f32(v)
        return self._intern(("const", struct.unpack("<I", struct.pack("<f", v))[0]))

    def const_value(self, i: int):
        k = self.nodes[i]
        if k[0] != "const":
            return None
        return struct.unpack("<f", struct.pack("<I", k[1]))[0]

    # operations (constant-folded when every input is a constant)
    def op(self, name: str, *args: int) -> int:
        vals = [self.const_value(a) for a in args]
        if all(v is not None for v in vals) and name in ("add", "sub", "mul", "div", "min",
                                                         "max", "neg", "sqrt"):
            return self.const(_apply(name, vals))
        return self._intern((name,) + tuple(args))

    def add(self, a, b):
        return self.op("add", a, b)

    def sub(self, a, b):
        return self.op("sub", a, b)

    def mul(self, a, b):
        return self.op("mul", a, b)

    def div(self, a, b):
        return self.op("div", a, b)

    def min(self, a, b):
        return self.op("min", a, b)

    def max(self, a, b):
        return self.op("max", a, b)

    def neg(self, a):
        return self.op("neg", a)

    def sqrt(self, a):
        return self.op("sqrt", a)

    def sin(self, a):
        return self.op("sin", a)

    def cos(self, a):
        return self.op("cos", a)

    def exp(self, a):
        return self.op("exp", a)

    def square(self, a):
        return self.mul(a, a)

    def abs(self, a):
        # The reference op set has no ABS: lib.cpp:68-71 uses sqrt(x*x).
        return self.sqrt(self.square(a))

    # evaluation of the DAG itself (used by tests to pin the tape compiler)
    def eval(self, root: int, x, y, z, t=0.0, fns=None):
        """Evaluate the DAG in float32 at one point; ``fns`` maps sin/cos/exp."""
        memo: dict[int, np.float32] = {}
        env = {"x": np.float32(x), "y": np.float32(y), "z": np.float32(z), "t": np.float32(t)}

        def rec(i):
            if i in memo:
                return memo[i]
            k = self.nodes[i]
            if k[0] == "axis":
                v = env[k[1]]
            elif k[0] == "const":
                v = np.float32(self.const_value(i))
            elif k[0] in ("sin", "cos", "exp"):
                v = np.float32(fns[k[0]](float(rec(k[1]))))
            else:
                v = np.float32(_apply(k[0], [rec(a) for a in k[1:]]))
            memo[i] = v
            return v

        order = self.topo(root)
        for i in order:
            rec(i)
        return float(memo[root])

    def topo(self, root: int) -> list[int]:
        """DFS post-order, inputs left to right (expression.cpp:298-314)."""
        seen, order = set(), []
        stack = [(root, False)]
        while stack:
            i, done = stack.pop()
            if done:
                order.append(i)
                continue
            if i in seen:
                continue
            seen.add(i)
            stack.append((i, True))
            k = self.nodes[i]
            if k[0] not in ("axis", "const"):
                for a in reversed(k[1:]):
                    if a not in seen:
                        stack.append((a, False))
        return order


@dataclass
class Tape:
    words: np.ndarray            # uint32 [n, instrs...]
    consts: np.ndarray           # float32 pool
    slot_count: int
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.words[0])

    def instrs(self):
        for w in self.words[1:1 + self.n]:
            w = int(w)
            yield (w & 0xFF, (w >> 8) & 0xFF, (w >> 16) & 0xFF, w >> 24)

    def op_histogram(self) -> dict[str, int]:
        h: dict[str, int] = {}
        for op, _, _, _ in self.instrs():
            h[OP_NAMES[op]] = h.get(OP_NAMES[op], 0) + 1
        return h

    def listing(self) -> str:
        lines = []
        for k, (op, o, a, b) in enumerate(self.instrs()):
            lines.append(f"{k:5d} {OP_NAMES[op]:>6} out={o} a={a} b={b}")
        return "\n".join(lines)


def encode(op: int, out: int, a: int = 0, b: int = 0) -> int:
    return (op & 0xFF) | ((out & 0xFF) << 8) | ((a & 0xFF) << 16) | ((b & 0xFF) << 24)


def compile_tape(bld: Builder, root: int) -> Tape:
    """Compile the DAG rooted at ``root`` (vulkan/src/csg/tape.cpp:14-228)."""
    order = bld.topo(root)
    pos = {node: k for k, node in enumerate(order)}
    # constant pool, first appearance (tape.cpp:44-55)
    pool: list[float] = []
    cidx: dict[int, int] = {}
    for node in order:
        key = bld.nodes[node]
        if key[0] == "const" and key[1] not in cidx:
            cidx[key[1]] = len(pool)
            pool.append(bld.const_value(node))
    if len(pool) > 0xFFFF:
        raise ValueError("constant pool exceeds 65536 entries")
    # liveness (tape.cpp:66-79)
    live = [-1] * len(order)
    for k, node in enumerate(order):
        key = bld.nodes[node]
        if key[0] not in ("axis", "const"):
            for a in key[1:]:
                live[pos[a]] = max(live[pos[a]], k)
    # slots 0..3 = x, y, z, t (tape.cpp:164-168)
    slots: list[int | None] = [None, None, None, None]
    for node in order:
        key = bld.nodes[node]
        if key[0] == "axis":
            slots[_AXES.index(key[1])] = node

    def free_slot():
        for s, v in enumerate(slots):
            if v is None:
                return s
        slots.append(None)
        if len(slots) > 256:
            raise ValueError("more than 256 slots")
        return len(slots) - 1

    def current(node):
        return slots.index(node)

    def release(node):
        if node in slots:
            slots[slots.index(node)] = None

    words: list[int] = []
    for k, node in enumerate(order):
        key = bld.nodes[node]
        kind = key[0]
        if kind == "axis":
            continue
        a = b = 0
        if kind == "const":
            ci = cidx[key[1]]
            a, b = ci & 0xFF, ci >> 8
            op = OP_CONST
        elif kind in _UNARY:
            src = key[1]
            a = current(src)
            if live[pos[src]] == k:
                release(src)
            op = _UNARY[kind]
        else:
            sa, sb = key[1], key[2]
            a, b = current(sa), current(sb)
            if live[pos[sa]] == k:
                release(sa)
            if live[pos[sb]] == k:
                release(sb)
            op = _BINARY[kind]
        out = free_slot()
        slots[out] = node
        words.append(encode(op, out, a, b))
    if bld.nodes[root][0] == "axis" and bld.nodes[root][1] != "x":
        # a bare y, z or t root: the result must still land in slot 0
        words.append(encode(OP_COPY, 0, _AXES.index(bld.nodes[root][1]), 0))
    if words and ((words[-1] >> 8) & 0xFF) != 0:
        raise AssertionError("result is not in slot 0")
    arr = np.array([len(words)] + words, dtype=np.uint32)
    return Tape(words=arr, consts=np.array(pool, dtype=np.float32),
                slot_count=max(4, len(slots)))


def tape_from_words(words, consts, slot_count) -> Tape:
    return Tape(words=np.asarray(words, dtype=np.uint32),
                consts=np.asarray(consts, dtype=np.float32), slot_count=int(slot_count))
