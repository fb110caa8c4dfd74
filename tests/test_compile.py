"""The tape compiler: frep_compile (libfrep.so, host code) against the oracle's
restatement of the reference compiler (oracle/csg_oracle.py), and the oracle
pinned by known-answer tapes derived by hand from vulkan/src/csg/
simplifier.cpp and tape.cpp, and by evaluation against the source DAG.

No device is used: frep_compile is host code.
"""
import numpy as np
import pytest

from frep import csg
from frep.csg import (Constant, Cos, Exp, Max, Min, Sin, Sqrt, X, Y, Z, T)
from frep.tape import encode

import csg_oracle as CO

CONST_, SIN_, COS_, ADD_, SUB_, MUL_, DIV_, MIN_, MAX_, EXP_, NEG_, SQRT_, COPY_ = range(13)


@pytest.fixture(scope="module")
def fns(oracle):
    return {"sin": oracle.sinf, "cos": oracle.cosf, "exp": oracle.expf}


def both(expr, fns):
    """(library tape, oracle words, oracle pool, oracle slots)"""
    nodes, root = csg.node_list(expr)
    tp = csg.compile_nodes(nodes, root)
    w, c, s = CO.compile_nodes(nodes, root, fns)
    return tp, w, c, s


def same(tp, w, c, s):
    return (np.array_equal(tp.words, w) and tp.slot_count == s
            and np.array_equal(np.asarray(tp.consts, np.float32).view(np.uint32),
                               np.asarray(c, np.float32).view(np.uint32)))


# ------------------------------------------------------------ known answers
# Each expected tape below is worked out by hand from the reference code:
# the fold rule that fires (simplifier.cpp:31-94), the operand order of
# NormalFormStep (ThreeWayCompare, OpOrder X<Y<Z<T<CONST<SIN<...<MAX, swap
# iff cmp(second, first) <= 0), then DFS order, constant pool, lowest free
# slot after releasing dead inputs (tape.cpp:151-228).
KATS = [
    # x * -1 -> NEG x (simplifier.cpp:62)
    ("mul_minus_one", lambda: X() * -1, [encode(NEG_, 0, 0)], []),
    # x / 4 -> x * 0.25 (simplifier.cpp:76): CONST takes slot 1 (y unused)
    ("div_const", lambda: X() / 4, [encode(CONST_, 1, 0), encode(MUL_, 0, 0, 1)], [0.25]),
    # -(y - x) -> x - y (simplifier.cpp:82)
    ("neg_sub", lambda: -(Y() - X()), [encode(SUB_, 0, 0, 1)], []),
    # y + x -> x + y (X sorts before Y)
    ("commute_axes", lambda: Y() + X(), [encode(ADD_, 0, 0, 1)], []),
    # 2 + x and x + 2 -> x + 2 (CONST sorts after X)
    ("commute_const_l", lambda: 2 + X(), [encode(CONST_, 1, 0), encode(ADD_, 0, 0, 1)], [2.0]),
    ("commute_const_r", lambda: X() + 2, [encode(CONST_, 1, 0), encode(ADD_, 0, 0, 1)], [2.0]),
    # min(x - 2, x - 1) -> min(x - 1, x - 2): SUB vs SUB, X == X, 1 < 2.
    # DFS: x, CONST 1 -> slot 1, SUB (CONST 1 dies first) -> slot 1,
    # CONST 2 -> slot 2, SUB (x and CONST 2 die) -> slot 0, MIN -> slot 0
    ("commute_structural", lambda: Min(X() - 2, X() - 1),
     [encode(CONST_, 1, 0), encode(SUB_, 1, 0, 1), encode(CONST_, 2, 1),
      encode(SUB_, 0, 0, 2), encode(MIN_, 0, 1, 0)], [1.0, 2.0]),
    # x + 0 -> x: bare x root is the empty tape
    ("add_zero", lambda: X() + 0, [], []),
    # 0 - y -> NEG y (simplifier.cpp:69)
    ("zero_minus", lambda: 0 - Y(), [encode(NEG_, 0, 1)], []),
    # x - (-y) -> x + y (simplifier.cpp:71)
    ("sub_neg", lambda: X() - (-Y()), [encode(ADD_, 0, 0, 1)], []),
    # x * 0 -> 0 (simplifier.cpp:58)
    ("mul_zero", lambda: Y() * 0, [encode(CONST_, 0, 0)], [0.0]),
    # -(-z) -> z: bare z root -> COPY 0 <- 2
    ("neg_neg", lambda: -(-Z()), [encode(COPY_, 0, 2)], []),
    # x / -1 -> NEG x (simplifier.cpp:75); 0 / y -> 0 (simplifier.cpp:73)
    ("div_minus_one", lambda: X() / -1, [encode(NEG_, 0, 0)], []),
    ("zero_div", lambda: 0 / Y(), [encode(CONST_, 0, 0)], [0.0]),
    # (-x) * y -> x * (-y) (simplifier.cpp:64-66); NEG y sorts after x
    ("mul_neg", lambda: (-X()) * Y(), [encode(NEG_, 1, 1), encode(MUL_, 0, 0, 1)], []),
    # (-x) + (-y): the reference returns -((-x) + (-y)) (simplifier.cpp:51-53)
    ("add_neg_neg", lambda: (-X()) + (-Y()),
     [encode(NEG_, 0, 0), encode(NEG_, 1, 1), encode(ADD_, 0, 0, 1), encode(NEG_, 0, 0)], []),
    # (-x) + y -> y - x (simplifier.cpp:54)
    ("add_neg_l", lambda: (-X()) + Y(), [encode(SUB_, 0, 1, 0)], []),
    # sqrt(x*x) twice, built separately, merged (simplifier.cpp:236-259)
    ("merge", lambda: Sqrt(X() * X()) + Sqrt(X() * X()),
     [encode(MUL_, 0, 0, 0), encode(SQRT_, 0, 0), encode(ADD_, 0, 0, 0)], []),
    # (x * y) + z -> z + (x * y) (Z sorts before MUL); DFS z, x, y, MUL:
    # x and y die at MUL -> slot 0; both die at ADD -> slot 0
    ("slots", lambda: X() * Y() + Z(), [encode(MUL_, 0, 0, 1), encode(ADD_, 0, 2, 0)], []),
    # sin(1) + x: all-constant op folded with the library's sin
    ("fold_sin", lambda: Sin(Constant(1.0)) + X(), None, None),
]


@pytest.mark.parametrize("name,build,words,pool", KATS, ids=[k[0] for k in KATS])
def test_oracle_known_answers(fns, name, build, words, pool):
    nodes, root = csg.node_list(build())
    w, c, s = CO.compile_nodes(nodes, root, fns)
    if name == "fold_sin":
        words = [encode(CONST_, 1, 0), encode(ADD_, 0, 0, 1)]
        pool = [fns["sin"](1.0)]
    assert list(w) == [len(words)] + words, name
    assert np.array_equal(np.asarray(c, np.float32), np.asarray(pool, np.float32))


def test_add_neg_neg_follows_the_reference(oracle, fns):
    # the reference's rewrite evaluates (-x) + (-y) as x + y
    nodes, root = csg.node_list((-X()) + (-Y()))
    w, c, s = CO.compile_nodes(nodes, root, fns)
    from frep.tape import tape_from_words
    tp = tape_from_words(w, c if len(c) else [0.0], s)
    assert oracle.eval_point(tp, 1.5, 2.0, 0.0) == 3.5


@pytest.mark.parametrize("name,build,words,pool", KATS, ids=[k[0] for k in KATS])
def test_library_matches_oracle_known_answers(fns, name, build, words, pool):
    assert same(*both(build(), fns))


# ------------------------------------------------------------ demo shapes
@pytest.mark.parametrize("name", sorted(csg.DEMO_SHAPES))
def test_demo_shapes_match_oracle(fns, name):
    tp, w, c, s = both(csg.DEMO_SHAPES[name](), fns)
    assert same(tp, w, c, s)
    assert tp.n > 10


# ------------------------------------------------------------ random DAGs
def random_dag(seed, n=60, allow_div=True, allow_neg_pairs=True):
    rng = np.random.default_rng(seed)
    pool = [X(), Y(), Z(), T()]
    consts = [0.0, 1.0, -1.0, 2.0, 0.5, -3.0, 0.25, 4.0, 1.5]
    for _ in range(n):
        k = rng.integers(0, 10)
        a = pool[rng.integers(0, len(pool))]
        b = pool[rng.integers(0, len(pool))]
        if rng.random() < 0.3:
            b = Constant(float(consts[rng.integers(0, len(consts))]))
        if rng.random() < 0.1:
            a = Constant(float(consts[rng.integers(0, len(consts))]))
        if k == 0:
            e = a + b
        elif k == 1:
            e = a - b
        elif k == 2:
            e = a * b
        elif k == 3:
            e = (a / b) if allow_div else (a * b)
        elif k == 4:
            e = Min(a, b)
        elif k == 5:
            e = Max(a, b)
        elif k == 6:
            e = -a
        elif k == 7:
            e = Sqrt(a * a + b * b)
        elif k == 8:
            e = Sin(a) if allow_div else Min(a, b)
        else:
            e = Exp(Min(a, Constant(1.0))) if allow_div else Max(a, b)
        pool.append(e)
    root = pool[-1]
    for e in pool[-8:-1]:
        root = Min(root, e) if rng.random() < 0.5 else Max(e, root)
    return root


@pytest.mark.parametrize("seed", range(40))
def test_random_dags_match_oracle(fns, seed):
    assert same(*both(random_dag(seed), fns))


@pytest.mark.parametrize("seed", range(20))
def test_compiled_tape_evaluates_like_the_dag(oracle, fns, seed):
    """Without division by constants (x / c becomes x * (1/c), which rounds
    differently) and transcendentals, every rewrite of the reference is
    value-preserving except for the sign of zero and (-a)+(-b), so the tape
    must give the DAG's value at every point."""
    rng = np.random.default_rng(100 + seed)
    e = random_dag(1000 + seed, n=40, allow_div=False)
    nodes, root = csg.node_list(e)
    w, c, s = CO.compile_nodes(nodes, root, fns)
    from frep.tape import tape_from_words
    tp = tape_from_words(w, c if len(c) else [0.0], s)
    folded_neg_pairs = False
    # the oracle's simplified DAG tells whether the (-a)+(-b) rule fired
    r = CO.merge_duplicates(CO.constant_fold(CO.from_nodes(nodes, root), fns))
    for x in CO.topo_iter(r):
        if x.op == CO.NEG and x.inputs[0].op == CO.ADD and \
                all(i.op == CO.NEG for i in x.inputs[0].inputs):
            folded_neg_pairs = True
    if folded_neg_pairs:
        pytest.skip("the (-a)+(-b) rewrite fired (tested on its own)")
    for p in rng.uniform(-2, 2, size=(50, 4)).astype(np.float32):
        want = e.eval(*p, fns=fns)
        got = oracle.eval_point(tp, *p)
        assert (np.isnan(want) and np.isnan(got)) or got == want, (p, got, want)


def test_compile_rejects_bad_node_lists():
    from frep import lib
    nodes = np.zeros(3, dtype=csg.NODE_DTYPE)
    nodes[0]["op"] = csg.X_
    nodes[1]["op"] = csg.ADD
    nodes[1]["a"], nodes[1]["b"] = 0, 2          # input after its user
    nodes[2]["op"] = csg.Y_
    with pytest.raises(lib.FrepError, match="earlier nodes"):
        csg.compile_nodes(nodes, 1)
    nodes[1]["op"] = 99
    with pytest.raises(lib.FrepError, match="unknown op"):
        csg.compile_nodes(nodes, 1)
    nodes = np.zeros(1, dtype=csg.NODE_DTYPE)
    nodes[0]["op"] = csg.CONST
    nodes[0]["value"] = np.nan
    with pytest.raises(lib.FrepError, match="NaN"):
        csg.compile_nodes(nodes, 0)


def test_wide_format_past_256_constants(fns):
    # MAX sorts after MIN, so each term follows the chain in DFS order and
    # few values are live at once; 300 distinct constants
    e = X()
    for k in range(300):
        e = Min(e, Max(X() - (float(k) * 0.37 + 1.0), Y()))
    tp, w, c, s = both(e, fns)
    assert same(tp, w, c, s)
    assert len(tp.consts) > 256 and tp.meta["format"] == 2 and tp.slot_count < 16


def test_more_than_256_live_values_rejected():
    # SUB sorts before MIN: every term is evaluated before the chain that
    # uses it, so 300 values are live at once (the reference asserts here)
    from frep import lib
    e = X()
    for k in range(300):
        e = Min(e, X() - (float(k) * 0.37 + 1.0))
    with pytest.raises(lib.FrepError, match="256 slots"):
        csg.compile(e)


def test_cpp_csg_api_builds_the_same_tapes():
    """include/frep_csg.hpp (the reference's csg API in C++, built into
    cpp/csg_demo by the library's Makefile) and frep/csg.py give the same
    tapes for the seven demo shapes."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "f-rep-voxel-rendering_amd", "csg_demo")
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    lines = {ln.split()[0]: ln.split() for ln in out.strip().splitlines()}
    assert sorted(lines) == sorted(csg.DEMO_SHAPES)
    for name, fn in csg.DEMO_SHAPES.items():
        tp = csg.compile(fn())
        _, n, slots, fmt, minmax, words, consts = lines[name]
        assert int(n) == tp.n and int(slots) == tp.slot_count and int(fmt) == tp.meta["format"]
        assert [int(w, 16) for w in words.split(",")] == [int(w) for w in tp.words]
        assert [int(c, 16) for c in consts.split(",")] == \
            [int(c) for c in np.asarray(tp.consts, np.float32).view(np.uint32)]
