"""CSG expressions, the reference's shape library and demo shapes, compiled to
tapes by frep_compile (libfrep.so).

Mirrors the reference's C++ API so shapes read the same:
  * Expr with +, -, *, / and unary - (vulkan/include/csg/expression.h:48-115,
    vulkan/src/csg/expression.cpp:79-107); numbers become constant nodes, one
    new node per use, as csg::Expr(float) does;
  * X, Y, Z, T, Constant, Sin, Cos, Exp, Sqrt, Min, Max
    (expression.cpp:225-290);
  * substitution e(x, y, z, t) (expression.cpp:281-296);
  * the library of vulkan/src/csg/lib.cpp:7-214 (Vector3, Square, Abs, Pow,
    Dot, Norm, Norm2, Dist, Dist2, Union, Intersect, Complement, Diff,
    Sphere, Box, Translate*, Scale*, Rotate*);
  * the demo shapes of vulkan/src/shapes.cpp:5-117.
Arithmetic the reference does on C++ floats (for example 1 / 3.0f in Menger)
is done here in float32.

compile(expr) hands the DAG to frep_compile as a node list and returns a
frep.tape.Tape with the compiler's words, constant pool and slot count.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .tape import Tape

X_, Y_, Z_, T_, CONST, SIN, COS, EXP, NEG, SQRT, ADD, SUB, MUL, DIV, MIN, MAX = range(16)
ARITY = [0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2]
NAMES = ["X", "Y", "Z", "T", "CONST", "SIN", "COS", "EXP", "NEG", "SQRT", "ADD", "SUB",
         "MUL", "DIV", "MIN", "MAX"]


def f32(v) -> float:
    return float(np.float32(v))


class Expr:
    """A node of the expression DAG; identity is the Python object."""
    __slots__ = ("op", "inputs", "value")

    def __init__(self, op: int, inputs=(), value: float = 0.0):
        self.op = op
        self.inputs = tuple(inputs)
        self.value = f32(value) if op == CONST else 0.0

    # arithmetic (expression.cpp:79-107)
    def __neg__(self):
        return Expr(NEG, (self,))

    def __add__(self, o):
        return Expr(ADD, (self, lift(o)))

    def __radd__(self, o):
        return Expr(ADD, (lift(o), self))

    def __sub__(self, o):
        return Expr(SUB, (self, lift(o)))

    def __rsub__(self, o):
        return Expr(SUB, (lift(o), self))

    def __mul__(self, o):
        return Expr(MUL, (self, lift(o)))

    def __rmul__(self, o):
        return Expr(MUL, (lift(o), self))

    def __truediv__(self, o):
        return Expr(DIV, (self, lift(o)))

    def __rtruediv__(self, o):
        return Expr(DIV, (lift(o), self))

    def __call__(self, nx, ny, nz, nt):
        """The DAG with x, y, z, t replaced (csg::Expr::operator(),
        expression.cpp:281-296): every operation node is rebuilt, constants
        are kept."""
        repl = {X_: lift(nx), Y_: lift(ny), Z_: lift(nz), T_: lift(nt)}
        out = {}
        for e in topo(self):
            if e.op in repl:
                out[id(e)] = repl[e.op]
            elif e.op == CONST:
                out[id(e)] = e
            else:
                out[id(e)] = Expr(e.op, tuple(out[id(i)] for i in e.inputs))
        return out[id(self)]

    def eval(self, x, y, z, t=0.0, fns=None) -> float:
        """Float32 evaluation of the DAG (no simplification); fns maps
        'sin', 'cos', 'exp' to the transcendentals to use."""
        F = np.float32
        env = {X_: F(x), Y_: F(y), Z_: F(z), T_: F(t)}
        val = {}
        with np.errstate(all="ignore"):
            for e in topo(self):
                if e.op in env:
                    v = env[e.op]
                elif e.op == CONST:
                    v = F(e.value)
                else:
                    a = val[id(e.inputs[0])]
                    b = val[id(e.inputs[1])] if len(e.inputs) > 1 else None
                    v = {SIN: lambda: F(fns["sin"](float(a))), COS: lambda: F(fns["cos"](float(a))),
                         EXP: lambda: F(fns["exp"](float(a))), NEG: lambda: -a,
                         SQRT: lambda: F(np.sqrt(a)), ADD: lambda: a + b, SUB: lambda: a - b,
                         MUL: lambda: a * b, DIV: lambda: a / b,
                         MIN: lambda: b if b < a else a, MAX: lambda: b if a < b else a}[e.op]()
                val[id(e)] = F(v)
        return float(val[id(self)])


def lift(v) -> Expr:
    return v if isinstance(v, Expr) else Expr(CONST, value=v)


def topo(root: Expr) -> list:
    """Depth-first post-order, inputs in order (expression.cpp:298-314)."""
    order, seen = [], {id(root)}
    stack = [(root, 0)]
    while stack:
        e, k = stack[-1]
        if k < len(e.inputs):
            stack[-1] = (e, k + 1)
            c = e.inputs[k]
            if id(c) not in seen:
                seen.add(id(c))
                stack.append((c, 0))
            continue
        order.append(e)
        stack.pop()
    return order


def X():
    return Expr(X_)


def Y():
    return Expr(Y_)


def Z():
    return Expr(Z_)


def T():
    return Expr(T_)


def Constant(v):
    return Expr(CONST, value=v)


def Sin(e):
    return Expr(SIN, (lift(e),))


def Cos(e):
    return Expr(COS, (lift(e),))


def Exp(e):
    return Expr(EXP, (lift(e),))


def Sqrt(e):
    return Expr(SQRT, (lift(e),))


def Min(a, b):
    return Expr(MIN, (lift(a), lift(b)))


def Max(a, b):
    return Expr(MAX, (lift(a), lift(b)))


# ------------------------------------------------------------ lib.cpp
class Vector3:
    __slots__ = ("x", "y", "z")

    def __init__(self, x, y, z):
        self.x, self.y, self.z = lift(x), lift(y), lift(z)

    def __add__(self, o):
        o = vec(o)
        return Vector3(self.x + o.x, self.y + o.y, self.z + o.z)

    def __sub__(self, o):
        o = vec(o)
        return Vector3(self.x - o.x, self.y - o.y, self.z - o.z)

    def __rsub__(self, o):
        return vec(o) - self

    def __mul__(self, o):
        if isinstance(o, Vector3) or isinstance(o, (tuple, list)):
            o = vec(o)
            return Vector3(self.x * o.x, self.y * o.y, self.z * o.z)
        return Vector3(self.x * o, self.y * o, self.z * o)

    def __rmul__(self, o):
        return Vector3(lift(o) * self.x, lift(o) * self.y, lift(o) * self.z)

    def __truediv__(self, o):
        if isinstance(o, Vector3) or isinstance(o, (tuple, list)):
            o = vec(o)
            return Vector3(self.x / o.x, self.y / o.y, self.z / o.z)
        return Vector3(self.x / o, self.y / o, self.z / o)

    def __neg__(self):
        return Vector3(-self.x, -self.y, -self.z)


def vec(v) -> Vector3:
    return v if isinstance(v, Vector3) else Vector3(*v)


def Empty():
    return Expr(CONST, value=1.0)


def Axes():
    return Vector3(X(), Y(), Z())


def Pi():
    return Expr(CONST, value=np.float32(np.pi))


def Square(x):
    x = lift(x)
    return x * x


def Abs(x):
    return Sqrt(Square(x))


def Pow(x, p: int):
    if p == 0:
        return Constant(1.0)
    if p == 1:
        return lift(x)
    if p % 2 == 0:
        return Square(Pow(x, p // 2))
    return Square(Pow(x, p // 2)) * x


def Dot(a, b):
    a, b = vec(a), vec(b)
    return a.x * b.x + a.y * b.y + a.z * b.z


def Norm(a):
    a = vec(a)
    return Sqrt(a.x * a.x + a.y * a.y + a.z * a.z)


def Norm2(a):
    a = vec(a)
    return a.x * a.x + a.y * a.y + a.z * a.z


def Dist(a, b):
    return Norm(vec(b) - vec(a))


def Dist2(a, b):
    return Norm2(vec(b) - vec(a))


def Union(a, b):
    return Min(a, b)


def Intersect(a, b):
    return Max(a, b)


def Complement(a):
    return -lift(a)


def Diff(a, b):
    return Intersect(a, Complement(b))


def Sphere(center, radius):
    return Dist2(Axes(), center) - Square(radius)


def Box(low, size):
    low, size = vec(low), vec(size)
    a = Square(X() - (low.x + size.x / 2)) - Square(size.x / 2)
    b = Square(Y() - (low.y + size.y / 2)) - Square(size.y / 2)
    c = Square(Z() - (low.z + size.z / 2)) - Square(size.z / 2)
    return Intersect(a, Intersect(b, c))


def TranslateX(a, dx):
    return a(X() - dx, Y(), Z(), T())


def TranslateY(a, dy):
    return a(X(), Y() - dy, Z(), T())


def TranslateZ(a, dz):
    return a(X(), Y(), Z() - dz, T())


def TranslateXYZ(a, d):
    d = vec(d)
    return a(X() - d.x, Y() - d.y, Z() - d.z, T())


def ScaleX(a, s):
    return a(X() / s, Y(), Z(), T())


def ScaleY(a, s):
    return a(X(), Y() / s, Z(), T())


def ScaleZ(a, s):
    return a(X(), Y(), Z() / s, T())


def ScaleXYZ(a, s):
    if isinstance(s, (Vector3, tuple, list)):
        s = vec(s)
        return a(X() / s.x, Y() / s.y, Z() / s.z, T())
    s = lift(s)
    return a(X() / s, Y() / s, Z() / s, T())


def RotateX(a, angle):
    c = Cos(-lift(angle))
    s = Sin(-lift(angle))
    return a(X(), c * Y() - s * Z(), s * Y() + c * Z(), T())


def RotateY(a, angle):
    c = Cos(-lift(angle))
    s = Sin(-lift(angle))
    return a(c * X() + s * Z(), Y(), -s * X() + c * Z(), T())


def RotateZ(a, angle):
    c = Cos(-lift(angle))
    s = Sin(-lift(angle))
    return a(c * X() - s * Y(), s * X() + c * Y(), Z(), T())


# ------------------------------------------------------------ shapes.cpp
def Screw():
    """A rotating screw (vulkan/src/shapes.cpp:5-19)."""
    shape = Box((-15, -2, -2), (30, 4, 4))
    shape = RotateX(shape, X())
    scale = 2 * Exp(X() / 40)
    shape = ScaleXYZ(shape, Vector3(1, scale, scale))
    shape = RotateZ(shape, T() / 2)
    shape = RotateY(shape, T() / 5)
    return shape


def ElasticCube():
    """A scaling cube (vulkan/src/shapes.cpp:21-36)."""
    shape = Box((-10, -10, -10), (20, 20, 20))
    scale = 1 / (1 + Pow(Y() / 5, 4))
    shape = ScaleX(shape, scale)
    shape = ScaleX(shape, 1 + f32(0.2) * Sin(Z() + 10 * T()))
    radius = (1 + f32(0.1) * Sin(T())) * 10
    shape = Diff(shape, Sphere((0, 0, 0), radius))
    shape = RotateY(shape, T() / 5)
    return shape


def TangleCube():
    """vulkan/src/shapes.cpp:38-45."""
    shape = Norm2(Axes() * Axes()) - 8 * Norm2(Axes()) + 25
    return ScaleXYZ(shape, 4)


def Morph():
    """vulkan/src/shapes.cpp:47-61."""
    shape1 = Norm2(Axes() * Axes()) - 8 * Norm2(Axes()) + 25
    shape1 = ScaleXYZ(shape1, 4)
    shape2 = Sphere((0, 0, 0), 10)
    blend = (1 + Sin(T())) / 2
    shape = blend * shape1 + (1 - blend) * shape2
    return RotateY(shape, T() / 5)


def Menger(level: int):
    """vulkan/src/shapes.cpp:63-81."""
    if level <= 0:
        return Box((0, 0, 0), (1, 1, 1))
    third = np.float32(1) / np.float32(3.0)
    block = ScaleXYZ(Menger(level - 1), Vector3(third, third, 1))
    menger = Empty()
    for x in range(3):
        for y in range(3):
            if not (x == 1 and y == 1):
                tb = TranslateXYZ(block, Vector3(np.float32(x) / np.float32(3.0),
                                                 np.float32(y) / np.float32(3.0), 0))
                menger = Union(menger, tb)
    return menger


def MengerSponge(level: int):
    """vulkan/src/shapes.cpp:83-91."""
    shape = Menger(level)
    shape = TranslateXYZ(shape, (-0.5, -0.5, -0.5))
    shape = ScaleXYZ(shape, Vector3(30, 30, 10))
    return Diff(shape, Sphere((10, 10, 0), 10))


def TwistedTower(level: int, angle: float, scale: float):
    """vulkan/src/shapes.cpp:93-106."""
    box = Box((-0.5, 0, -0.5), (1, 1, 1))
    if level <= 0:
        return box
    shape = TwistedTower(level - 1, angle, scale)
    shape = ScaleXYZ(shape, scale)
    shape = RotateY(shape, angle)
    shape = TranslateY(shape, scale)
    return Union(shape, box)


def OrganicBalls():
    """vulkan/src/shapes.cpp:108-117."""
    s1 = Sphere((10, 0, 0), 8)
    s1 = TranslateX(s1, 5 * Sin(T()))
    s2 = Sphere((-10, 0, 0), 10)
    f = Min(Min(X(), Y()), Sqrt(X() * Y()) - 40)
    return f(s1, s2, Z(), T())


# The seven demo shapes with the parameters used here (MengerSponge's level
# 3 is the commented-out choice of vulkan/src/main.cpp:18; TwistedTower has
# no parameters in the reference, these are this build's).
DEMO_SHAPES = {
    "screw": Screw,
    "elastic_cube": ElasticCube,
    "tangle_cube": TangleCube,
    "morph": Morph,
    "menger_sponge3": lambda: MengerSponge(3),
    "twisted_tower": lambda: TwistedTower(6, 0.35, 0.8),
    "organic_balls": OrganicBalls,
}
# The reference's world box for them (vulkan/src/application.cpp:21-22).
DEMO_LOW = (-20.0, -20.0, -20.0)
DEMO_WORLD = 40.0


# ------------------------------------------------------------ compile
NODE_DTYPE = np.dtype([("op", "<u4"), ("a", "<u4"), ("b", "<u4"), ("value", "<f4")])


def node_list(root: Expr):
    """The DAG as frep_expr_node records (inputs before users) and the
    root's index."""
    order = topo(root)
    idx = {id(e): k for k, e in enumerate(order)}
    arr = np.zeros(len(order), dtype=NODE_DTYPE)
    for k, e in enumerate(order):
        arr[k]["op"] = e.op
        if e.inputs:
            arr[k]["a"] = idx[id(e.inputs[0])]
        if len(e.inputs) > 1:
            arr[k]["b"] = idx[id(e.inputs[1])]
        if e.op == CONST:
            arr[k]["value"] = e.value
    return arr, len(order) - 1


class CompileInfo(C.Structure):
    _fields_ = [("n_words", C.c_uint32), ("n_consts", C.c_uint32), ("slot_count", C.c_uint32),
                ("format", C.c_uint32), ("dag_nodes", C.c_uint32), ("tape_nodes", C.c_uint32)]


def compile_nodes(nodes: np.ndarray, root: int) -> Tape:
    """frep_compile on a node list."""
    from . import lib
    L = lib.load()
    nodes = np.ascontiguousarray(nodes, dtype=NODE_DTYPE)
    info = CompileInfo()
    np_ = nodes.ctypes.data_as(C.c_void_p)
    st = L.frep_compile(np_, len(nodes), root, None, 0, None, 0, C.byref(info))
    if st != 0:
        raise lib.FrepError(st, L.frep_compile_last_error().decode())
    words = np.zeros(info.n_words, dtype=np.uint32)
    consts = np.zeros(max(1, info.n_consts), dtype=np.float32)
    st = L.frep_compile(np_, len(nodes), root, words.ctypes.data_as(C.POINTER(C.c_uint32)),
                        len(words), consts.ctypes.data_as(C.POINTER(C.c_float)), len(consts),
                        C.byref(info))
    if st != 0:
        raise lib.FrepError(st, L.frep_compile_last_error().decode())
    return Tape(words=words, consts=consts[:info.n_consts], slot_count=int(info.slot_count),
                meta={"format": int(info.format), "dag_nodes": int(info.dag_nodes),
                      "tape_nodes": int(info.tape_nodes)})


def compile(expr: Expr) -> Tape:
    """csg::Tape(expr) (vulkan/src/csg/tape.cpp:14-21) through frep_compile."""
    nodes, root = node_list(expr)
    return compile_nodes(nodes, root)
